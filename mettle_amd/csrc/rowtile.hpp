// Row-tile helpers shared by the scan kernels (filter, semi-join, fact
// pass): 16 rows per lane as 4 steps of 4 consecutive rows, so each load
// instruction of a wave covers one contiguous span (16-byte lanes), typed
// predicate dispatch, and key loads.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.hpp"

namespace mtl {
namespace dev {

// Tile layout: 4 waves x FJ steps x 64 lanes x 4 consecutive rows, so lane l
// of wave w at step j owns rows base + w*FJ*256 + j*256 + 4l + {0..3}: every
// load instruction of a wave covers one contiguous span (16-byte lanes).
// ---------------------------------------------------------------------------
constexpr int FBLOCK = 256;
constexpr int FJ = 4;
constexpr int FTILE = FBLOCK * FJ * 4;  // 4096 rows per tile
constexpr int FWAVES = FBLOCK / WAVE;

typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned long long u64x2_t __attribute__((ext_vector_type(2)));

// Streaming loads: a column a scan reads once goes through with the
// nontemporal hint, so it does not displace cached data (bitmaps, tables)
// and the memory system treats it as a stream.  Measured on MI355X for a
// two-column 480 MB tile scan: 6.3 TB/s plain, 7.0 TB/s nontemporal.
__device__ __forceinline__ uint4 ld_stream(const void* p) {
  const u32x4_t t = __builtin_nontemporal_load((const u32x4_t*)p);
  return make_uint4(t.x, t.y, t.z, t.w);
}
__device__ __forceinline__ ulonglong2 ld_stream2(const void* p) {
  const u64x2_t t = __builtin_nontemporal_load((const u64x2_t*)p);
  ulonglong2 r;
  r.x = t.x;
  r.y = t.y;
  return r;
}
__device__ __forceinline__ uint32_t ld_stream1(const void* p) {
  return __builtin_nontemporal_load((const uint32_t*)p);
}

// FULL: the caller knows the four rows are in range and the column is
// aligned, so the load is one straight-line vector load, with the streaming
// hint unless NT is false (a tile several blocks read, which should stay in
// L2).  (A branch around each load makes the compiler wait for it before
// the next is issued.)
template <typename T, bool FULL = false, bool NT = true>
__device__ __forceinline__ void load4(const T* __restrict__ col, uint64_t k,
                                      uint64_t n, bool vec_ok, T x[4]) {
  constexpr bool S = FULL && NT;
  if (FULL || (vec_ok && k + 3 < n)) {
    if constexpr (sizeof(T) == 4) {
      uint4 u = S ? ld_stream(col + k) : *(const uint4*)(col + k);
      x[0] = __builtin_bit_cast(T, u.x);
      x[1] = __builtin_bit_cast(T, u.y);
      x[2] = __builtin_bit_cast(T, u.z);
      x[3] = __builtin_bit_cast(T, u.w);
    } else if constexpr (sizeof(T) == 8) {
      ulonglong2 a = S ? ld_stream2(col + k) : *(const ulonglong2*)(col + k);
      ulonglong2 b = S ? ld_stream2(col + k + 2) : *(const ulonglong2*)(col + k + 2);
      x[0] = __builtin_bit_cast(T, a.x);
      x[1] = __builtin_bit_cast(T, a.y);
      x[2] = __builtin_bit_cast(T, b.x);
      x[3] = __builtin_bit_cast(T, b.y);
    } else {
      uint32_t u = S ? ld_stream1(col + k) : *(const uint32_t*)(col + k);
      x[0] = (T)(u & 0xff);
      x[1] = (T)((u >> 8) & 0xff);
      x[2] = (T)((u >> 16) & 0xff);
      x[3] = (T)(u >> 24);
    }
  } else {
#pragma unroll
    for (int v = 0; v < 4; ++v) x[v] = (k + v < n) ? col[k + v] : T(0);
  }
}

// Raw bytes of 4 consecutive rows (zero-extended): one 16-byte load (two
// for 8-byte values, one 4-byte load for bytes) when aligned and in range.
__device__ __forceinline__ void load_raw4(const void* col, uint32_t width, uint64_t k,
                                          uint64_t n, bool vec_ok, uint64_t (&x)[4]) {
  if (vec_ok && k + 3 < n) {
    if (width == 4) {
      uint4 u = *(const uint4*)((const uint32_t*)col + k);
      x[0] = u.x; x[1] = u.y; x[2] = u.z; x[3] = u.w;
    } else if (width == 8) {
      ulonglong2 a = *(const ulonglong2*)((const uint64_t*)col + k);
      ulonglong2 b = *(const ulonglong2*)((const uint64_t*)col + k + 2);
      x[0] = a.x; x[1] = a.y; x[2] = b.x; x[3] = b.y;
    } else {
      uint32_t u = *(const uint32_t*)((const uint8_t*)col + k);
      x[0] = u & 0xFFu; x[1] = (u >> 8) & 0xFFu; x[2] = (u >> 16) & 0xFFu; x[3] = u >> 24;
    }
  } else {
#pragma unroll
    for (int v = 0; v < 4; ++v) x[v] = k + v < n ? load_raw(col, width, k + v) : 0ull;
  }
}

// Dense rows k0 + j*256 + v: clear the bits of rows failing p.  All four
// steps are loaded before any compare: a "skip dead step" branch here would
// depend on the bits the previous step just updated and serialise the loads.
__device__ __forceinline__ void pred_dense(const PredC& p, uint64_t k0, uint64_t n, bool vec_ok,
                                           uint32_t& bits) {
  const uint32_t width = pred_width(p.kind);
  uint64_t x[FJ][4];
#pragma unroll
  for (int j = 0; j < FJ; ++j) load_raw4(p.col, width, k0 + (uint64_t)j * 256, n, vec_ok, x[j]);
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (!pred_pass(enc_raw(x[j][v], p.kind), p)) bits &= ~(1u << (4 * j + v));
}

// Rows given by row id (all loads first, then the compares).
__device__ __forceinline__ void pred_sel(const PredC& p, const uint32_t (&rid)[FJ][4],
                                         uint32_t& bits) {
  const uint32_t width = pred_width(p.kind);
  uint64_t x[FJ][4];
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      x[j][v] = (bits >> (4 * j + v)) & 1u ? load_raw(p.col, width, rid[j][v]) : 0ull;
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      uint32_t b = 1u << (4 * j + v);
      if ((bits & b) && !pred_pass(enc_raw(x[j][v], p.kind), p)) bits &= ~b;
    }
}

// pred_sel for dense rows k0 + j*256 + v (every row with a set bit is < n),
// one step of four rows at a time: a later predicate of a scan that has
// already dropped most rows, kept small in registers.
__device__ __forceinline__ void pred_rows(const PredC& p, uint64_t k0, uint32_t& bits) {
  const uint32_t width = pred_width(p.kind);
#pragma unroll
  for (int j = 0; j < FJ; ++j) {
    const uint32_t m = (bits >> (4 * j)) & 0xFu;
    if (!m) continue;
    uint64_t x[4];
#pragma unroll
    for (int v = 0; v < 4; ++v)
      x[v] = (m >> v) & 1u ? load_raw(p.col, width, k0 + (uint64_t)j * 256 + v) : 0ull;
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (((m >> v) & 1u) && !pred_pass(enc_raw(x[v], p.kind), p))
        bits &= ~(1u << (4 * j + v));
  }
}

// Integer keys of the lane's 16 rows: 16-byte vector loads when the rows
// are dense (no input selection), else one load per row id (row ids of
// rows past the end are 0, so every load is in bounds).
__device__ __forceinline__ void load_keys16(const void* col, int32_t type, bool dense,
                                            uint64_t k0, uint64_t n,
                                            const uint32_t (&rid)[FJ][4],
                                            int64_t (&key)[FJ][4]) {
  if (dense) {
#pragma unroll
    for (int j = 0; j < FJ; ++j) {
      uint64_t k = k0 + (uint64_t)j * 256;
      switch (type) {
        case MTL_BIGINT:
        case MTL_TIMESTAMP:
        case MTL_DECIMAL: {
          int64_t x[4];
          load4<int64_t>((const int64_t*)col, k, n, true, x);
#pragma unroll
          for (int v = 0; v < 4; ++v) key[j][v] = x[v];
          break;
        }
        case MTL_DICT8: {
          uint8_t x[4];
          load4<uint8_t>((const uint8_t*)col, k, n, true, x);
#pragma unroll
          for (int v = 0; v < 4; ++v) key[j][v] = x[v];
          break;
        }
        default: {
          int32_t x[4];
          load4<int32_t>((const int32_t*)col, k, n, true, x);
#pragma unroll
          for (int v = 0; v < 4; ++v) key[j][v] = x[v];
          break;
        }
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) key[j][v] = load_key(col, type, rid[j][v]);
  }
}

// Raw values (zero-extended) of the 16 dense rows of a lane; 0 past n.
__device__ __forceinline__ void load_u32x16(const void* col, uint32_t width, uint64_t k0,
                                            uint64_t n, bool vec, uint32_t (&x)[FJ][4]) {
#pragma unroll
  for (int j = 0; j < FJ; ++j) {
    const uint64_t k = k0 + (uint64_t)j * 256;
    if (vec && k + 3 < n) {
      if (width == 4) {
        uint4 u = *(const uint4*)((const uint32_t*)col + k);
        x[j][0] = u.x; x[j][1] = u.y; x[j][2] = u.z; x[j][3] = u.w;
      } else {
        uint32_t u = *(const uint32_t*)((const uint8_t*)col + k);
        x[j][0] = u & 0xFFu; x[j][1] = (u >> 8) & 0xFFu; x[j][2] = (u >> 16) & 0xFFu;
        x[j][3] = u >> 24;
      }
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v)
        x[j][v] = k + v < n ? (width == 4 ? ((const uint32_t*)col)[k + v]
                                          : (uint32_t)((const uint8_t*)col)[k + v])
                            : 0u;
    }
  }
}

// Bits of the lane's 16 rows that lie below n.
__device__ __forceinline__ uint32_t rows_below(uint64_t k0, uint64_t n) {
  uint32_t bits = 0;
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (k0 + (uint64_t)j * 256 + v < n) bits |= 1u << (4 * j + v);
  return bits;
}

// An empty asm statement the compiler must assume rewrites x.  Called at
// the top of a loop over queries on a lane's 16 loaded rows, it keeps the
// work derived from those rows inside the loop: hoisted out, the per-row
// values (encodings, masks, addresses) would all stay live across the loop
// and drive the kernel to the register limit.
template <typename T>
__device__ __forceinline__ void keep_in_loop(T& x) {
  asm volatile("" : "+v"(x));
}
template <typename T>
__device__ __forceinline__ void keep_in_loop(T (&x)[FJ][4]) {
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) keep_in_loop(x[j][v]);
}

// Raw narrow key -> offset from min, 0xFFFFFFFF when outside [0, span).
__device__ __forceinline__ uint32_t narrow_off(uint32_t raw, int32_t type, int64_t min,
                                               uint64_t span) {
  const int64_t key = type == MTL_DICT8 ? (int64_t)raw : (int64_t)(int32_t)raw;
  const uint64_t o = key_off(key, min);
  return o < span ? (uint32_t)o : 0xFFFFFFFFu;
}

// The pass bits of one 1024-row wave segment: lane l's 4 bits of step j sit
// at bits 4(l%16).. of word l/16 of the step's 256 rows.  An OR scan over
// each 16-lane row (DPP row shifts) leaves the word in the row's last lane,
// which stores it.
__device__ __forceinline__ void store_pass_bits(uint64_t* sel_bm, uint64_t seg, uint32_t bits,
                                                int lane) {
  const uint32_t r = (uint32_t)lane & 15u;
#pragma unroll
  for (int j = 0; j < FJ; ++j) {
    const uint32_t nib = (bits >> (4 * j)) & 0xFu;
    uint32_t lo = r < 8 ? nib << (4 * r) : 0u;
    uint32_t hi = r < 8 ? 0u : nib << (4 * (r - 8));
    lo |= row_shr<1>(lo); hi |= row_shr<1>(hi);
    lo |= row_shr<2>(lo); hi |= row_shr<2>(hi);
    lo |= row_shr<4>(lo); hi |= row_shr<4>(hi);
    lo |= row_shr<8>(lo); hi |= row_shr<8>(hi);
    if (r == 15) sel_bm[(seg + (uint64_t)j * 256) / 64 + (lane >> 4)] = ((uint64_t)hi << 32) | lo;
  }
}

constexpr uint32_t kMQQueries = 64;  // queries per shared scan pass

}  // namespace dev
}  // namespace mtl
