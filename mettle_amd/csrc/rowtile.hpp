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

template <typename T>
__device__ __forceinline__ void load4(const T* __restrict__ col, uint64_t k,
                                      uint64_t n, bool vec_ok, T x[4]) {
  if (vec_ok && k + 3 < n) {
    if constexpr (sizeof(T) == 4) {
      uint4 u = *(const uint4*)(col + k);
      x[0] = __builtin_bit_cast(T, u.x);
      x[1] = __builtin_bit_cast(T, u.y);
      x[2] = __builtin_bit_cast(T, u.z);
      x[3] = __builtin_bit_cast(T, u.w);
    } else if constexpr (sizeof(T) == 8) {
      ulonglong2 a = *(const ulonglong2*)(col + k);
      ulonglong2 b = *(const ulonglong2*)(col + k + 2);
      x[0] = __builtin_bit_cast(T, a.x);
      x[1] = __builtin_bit_cast(T, a.y);
      x[2] = __builtin_bit_cast(T, b.x);
      x[3] = __builtin_bit_cast(T, b.y);
    } else {
      uint32_t u = *(const uint32_t*)(col + k);
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

template <typename T, typename Cmp>
__device__ __forceinline__ void pred_dense(const T* __restrict__ col,
                                           uint64_t k0, uint64_t n, bool vec_ok,
                                           Cmp cmp, uint32_t& bits) {
#pragma unroll
  for (int j = 0; j < FJ; ++j) {
    uint32_t m = (bits >> (4 * j)) & 0xFu;
    if (!m) continue;
    T x[4];
    load4<T>(col, k0 + (uint64_t)j * 256, n, vec_ok, x);
    uint32_t r = 0;
#pragma unroll
    for (int v = 0; v < 4; ++v) r |= (uint32_t)cmp(x[v]) << v;
    bits &= ~(((~r) & 0xFu) << (4 * j));
  }
}

template <typename T, typename Cmp>
__device__ __forceinline__ void pred_sel(const T* __restrict__ col,
                                         const uint32_t (&rid)[FJ][4],
                                         Cmp cmp, uint32_t& bits) {
  // all loads first (they overlap), then the compares
  T x[FJ][4];
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      x[j][v] = (bits >> (4 * j + v)) & 1u ? col[rid[j][v]] : T(0);
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      uint32_t b = 1u << (4 * j + v);
      if ((bits & b) && !cmp(x[j][v])) bits &= ~b;
    }
}

// pred_sel for dense rows: row of (j, v) is k0 + j*256 + v (every row with
// a set bit is below n).
template <typename T, typename Cmp>
__device__ __forceinline__ void pred_rows(const T* __restrict__ col, uint64_t k0, Cmp cmp,
                                          uint32_t& bits) {
  T x[FJ][4];
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      x[j][v] = (bits >> (4 * j + v)) & 1u ? col[k0 + (uint64_t)j * 256 + v] : T(0);
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      uint32_t b = 1u << (4 * j + v);
      if ((bits & b) && !cmp(x[j][v])) bits &= ~b;
    }
}

// Dispatch one predicate on (type, op); `f` receives a typed column pointer
// and a comparison functor.
template <typename F>
__device__ __forceinline__ void with_pred(const mtl_pred& p, F&& f) {
  const int op = p.op;
  switch (p.type) {
    case MTL_INTEGER:
    case MTL_DATE:
    case MTL_DICT32: {
      const int32_t s = (int32_t)p.scalar;
      const int32_t* c = (const int32_t*)p.column;
      switch (op) {
        case MTL_CMP_EQ: f(c, [=](int32_t x) { return x == s; }); break;
        case MTL_CMP_LT: f(c, [=](int32_t x) { return x < s; }); break;
        case MTL_CMP_LE: f(c, [=](int32_t x) { return x <= s; }); break;
        case MTL_CMP_GT: f(c, [=](int32_t x) { return x > s; }); break;
        default: f(c, [=](int32_t x) { return x >= s; }); break;
      }
      break;
    }
    case MTL_BIGINT:
    case MTL_TIMESTAMP:
    case MTL_DECIMAL: {
      const int64_t s = p.scalar;
      const int64_t* c = (const int64_t*)p.column;
      switch (op) {
        case MTL_CMP_EQ: f(c, [=](int64_t x) { return x == s; }); break;
        case MTL_CMP_LT: f(c, [=](int64_t x) { return x < s; }); break;
        case MTL_CMP_LE: f(c, [=](int64_t x) { return x <= s; }); break;
        case MTL_CMP_GT: f(c, [=](int64_t x) { return x > s; }); break;
        default: f(c, [=](int64_t x) { return x >= s; }); break;
      }
      break;
    }
    case MTL_DICT8: {
      const int32_t s = (int32_t)p.scalar;
      const uint8_t* c = (const uint8_t*)p.column;
      switch (op) {
        case MTL_CMP_EQ: f(c, [=](uint8_t x) { return (int32_t)x == s; }); break;
        case MTL_CMP_LT: f(c, [=](uint8_t x) { return (int32_t)x < s; }); break;
        case MTL_CMP_LE: f(c, [=](uint8_t x) { return (int32_t)x <= s; }); break;
        case MTL_CMP_GT: f(c, [=](uint8_t x) { return (int32_t)x > s; }); break;
        default: f(c, [=](uint8_t x) { return (int32_t)x >= s; }); break;
      }
      break;
    }
    case MTL_FLOAT: {
      const float s = (float)__longlong_as_double(p.scalar);
      const float* c = (const float*)p.column;
      f(c, [=](float x) { return cmp_float<float>(x, s, op); });
      break;
    }
    case MTL_DOUBLE: {
      const double s = __longlong_as_double(p.scalar);
      const double* c = (const double*)p.column;
      f(c, [=](double x) { return cmp_float<double>(x, s, op); });
      break;
    }
    default:
      break;
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

}  // namespace dev
}  // namespace mtl
