// Predicate evaluation, stable selection-vector compaction, gather and
// scalar projection.
//
// Reference semantics: compare_{i32,i64,f32}_scalar (kernels.metal:70-173),
// and_masks (:673-682), filter_to_indices (:684-696), gather_*_v
// (:725-767), project_*_scalar (:4-68, 958-975).  Differences by design:
// the conjunction, mask and compaction are one pass (no u8 mask in HBM);
// compaction is stable (ballot/popcount ranks + decoupled look-back) where
// the reference used an unordered atomic counter; projections are exact
// integer arithmetic with overflow detection instead of float32.
#include "device.hpp"
#include "filter.hpp"
#include "mtl_internal.hpp"
#include "prof.hpp"

namespace mtl {
using namespace dev;

// ---------------------------------------------------------------------------
// Clearing kernel: zeroes up to 8 regions in one launch (used instead of
// hipMemsetAsync so that the sequence is safe inside captured graphs).
// ---------------------------------------------------------------------------
__global__ void k_zero(ZeroList zl) {
  for (int r = 0; r < zl.n; ++r) {
    uint32_t* p = zl.ptr[r];
    uint64_t words = zl.words[r];
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
         i += (uint64_t)gridDim.x * blockDim.x)
      p[i] = zl.value[r];
  }
}

int32_t launch_zero(const ZeroList& zl, hipStream_t st) {
  uint64_t mx = 0;
  for (int r = 0; r < zl.n; ++r) mx = zl.words[r] > mx ? zl.words[r] : mx;
  if (zl.n == 0 || mx == 0) return MTL_OK;
  uint32_t blocks = (uint32_t)std::min<uint64_t>(ceil_div(mx, 256), 2048);
  hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(256), 0, st, zl);
  MTL_HIP(hipGetLastError());
  return MTL_OK;
}

// ---------------------------------------------------------------------------
// Filter + compaction.
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
#pragma unroll
  for (int j = 0; j < FJ; ++j) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      uint32_t b = 1u << (4 * j + v);
      if ((bits & b) && !cmp(col[rid[j][v]])) bits &= ~b;
    }
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

// Evaluate all predicates (and the optional semi-join probe) for the 16 rows
// of this lane; returns the pass bits (bit 4j+v).
__device__ __forceinline__ uint32_t eval_lane(const FilterArgs& a, uint64_t n,
                                              uint64_t k0,
                                              uint32_t (&rid)[FJ][4]) {
  uint32_t bits = 0;
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (k0 + (uint64_t)j * 256 + v < n) bits |= 1u << (4 * j + v);
  if (a.in_sel) {
#pragma unroll
    for (int j = 0; j < FJ; ++j) {
      uint64_t k = k0 + (uint64_t)j * 256;
      if (k + 3 < n) {
        uint4 u = *(const uint4*)(a.in_sel + k);
        rid[j][0] = u.x; rid[j][1] = u.y; rid[j][2] = u.z; rid[j][3] = u.w;
      } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) rid[j][v] = (k + v < n) ? a.in_sel[k + v] : 0u;
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) rid[j][v] = (uint32_t)(k0 + (uint64_t)j * 256 + v);
  }

  for (int p = 0; p < a.n_preds && bits; ++p) {
    const bool vec_ok = (a.vec_ok >> p) & 1u;
    if (a.in_sel) {
      with_pred(a.preds[p], [&](auto col, auto cmp) { pred_sel(col, rid, cmp, bits); });
    } else {
      with_pred(a.preds[p], [&](auto col, auto cmp) {
        pred_dense(col, k0, n, vec_ok, cmp, bits);
      });
    }
  }
  // Semi-join probe: key(row) - min must be a set bit of the bitmap.
  if (a.probe_bm && bits) {
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        uint32_t b = 1u << (4 * j + v);
        if (bits & b) {
          int64_t key = load_key(a.probe_key, a.probe_key_type, rid[j][v]);
          uint64_t s = key_off(key, a.probe_min);
          bool hit = s < a.probe_span &&
                     ((a.probe_bm[s >> 6] >> (s & 63)) & 1ull);
          if (!hit) bits &= ~b;
        }
      }
  }
  return bits;
}

__global__ __launch_bounds__(FBLOCK) void k_filter(FilterArgs a) {
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_wave_tot[FWAVES];
  __shared__ uint32_t s_prefix;
  const int tid = threadIdx.x;
  const int wave = tid / WAVE;
  const int lane = tid % WAVE;
  if (tid == 0) s_tile = atomicAdd(a.tile_ctr, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t n = a.in_sel ? (a.in_count ? (uint64_t)*a.in_count : a.n_rows)
                              : a.n_rows;
  const uint64_t k0 = (uint64_t)tile * FTILE + (uint64_t)wave * (FJ * 256) +
                      (uint64_t)lane * 4;
  uint32_t rid[FJ][4];
  uint32_t bits = 0;
  if ((uint64_t)tile * FTILE < n) bits = eval_lane(a, n, k0, rid);
  else {
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) rid[j][v] = 0;
  }

  // per-step ballots -> ranks in row order
  uint32_t lane_base[FJ];
  uint32_t wave_total = 0;
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int j = 0; j < FJ; ++j) {
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
  if (lane == 0) s_wave_tot[wave] = wave_total;
  __syncthreads();
  uint32_t wave_off = 0, tile_total = 0;
#pragma unroll
  for (int w = 0; w < FWAVES; ++w) {
    uint32_t t = s_wave_tot[w];
    if (w < wave) wave_off += t;
    tile_total += t;
  }
  if (wave == 0) {
    uint32_t ex = lookback_wave(a.status, tile, tile_total, a.err);
    if (lane == 0) s_prefix = ex;
  }
  __syncthreads();
  const uint32_t base = s_prefix + wave_off;
  if (bits) {
#pragma unroll
    for (int j = 0; j < FJ; ++j) {
      uint32_t pos = base + lane_base[j];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        if ((bits >> (4 * j + v)) & 1u) {
          uint32_t r = rid[j][v];
          a.out_sel[pos] = r;
          if (a.mark_bm) {
            int64_t key = load_key(a.mark_key, a.mark_key_type, r);
            uint64_t s = key_off(key, a.mark_min);
            if (s < a.mark_span) {
              unsigned long long bit = 1ull << (s & 63);
              unsigned long long old =
                  atomicOr((unsigned long long*)&a.mark_bm[s >> 6], bit);
              if (old & bit) atomicOr(a.err, 2u);  // duplicate key
            } else {
              atomicOr(a.err, 8u);  // key outside the stated [min, max]
            }
          }
          ++pos;
        }
      }
    }
  }
  if (tile == gridDim.x - 1 && tid == 0) *a.out_count = s_prefix + tile_total;
}

uint32_t filter_tiles(uint64_t n) { return (uint32_t)ceil_div(n ? n : 1, FTILE); }

int32_t launch_filter(FilterArgs a, uint64_t capacity, hipStream_t st) {
  uint32_t tiles = filter_tiles(capacity);
  a.vec_ok = 0;
  for (uint32_t p = 0; p < a.n_preds; ++p) {
    int w = type_width(a.preds[p].type);
    uintptr_t addr = (uintptr_t)a.preds[p].column;
    int need = w == 1 ? 4 : 16;
    if (addr % need == 0) a.vec_ok |= 1u << p;
  }
  if (a.in_sel && ((uintptr_t)a.in_sel % 16) != 0)
    return set_error(MTL_ERR_INVALID_ARGUMENT,
                     "selection vectors must be 16-byte aligned");
  hipLaunchKernelGGL(k_filter, dim3(tiles), dim3(FBLOCK), 0, st, a);
  MTL_HIP(hipGetLastError());
  return MTL_OK;
}

// ---------------------------------------------------------------------------
// Mask output (compare_*_values counterpart).
// ---------------------------------------------------------------------------
struct CompareArgs {
  mtl_pred preds[MTL_MAX_PREDS];
  uint32_t n_preds;
  uint64_t n;
  uint8_t* out;
};

__global__ void k_compare(CompareArgs a) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    bool ok = true;
    for (uint32_t p = 0; p < a.n_preds && ok; ++p) ok = eval_pred_row(a.preds[p], i);
    a.out[i] = ok ? 1 : 0;
  }
}

// ---------------------------------------------------------------------------
// Gather through a selection (count on the device).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_gather(const T* __restrict__ col, const uint32_t* __restrict__ sel,
                         const uint32_t* count, uint64_t cap, T* __restrict__ out) {
  const uint64_t n = count ? (uint64_t)*count : cap;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = col[sel ? sel[i] : i];
}

// ---------------------------------------------------------------------------
// Scalar projection with DuckDB overflow semantics.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T proj_int(T x, T s, int op, bool& ovf) {
  T r;
  bool o;
  switch (op) {
    case MTL_PROJ_ADD: o = __builtin_add_overflow(x, s, &r); break;
    case MTL_PROJ_MUL: o = __builtin_mul_overflow(x, s, &r); break;
    case MTL_PROJ_SUB: o = __builtin_sub_overflow(x, s, &r); break;
    default: o = __builtin_sub_overflow(s, x, &r); break;
  }
  ovf |= o;
  return r;
}

template <typename T>
__device__ __forceinline__ T proj_fp(T x, T s, int op) {
  switch (op) {
    case MTL_PROJ_ADD: return x + s;
    case MTL_PROJ_MUL: return x * s;
    case MTL_PROJ_SUB: return x - s;
    default: return s - x;
  }
}

template <typename T, bool IS_FP>
__global__ void k_project(const T* __restrict__ col, const uint32_t* __restrict__ sel,
                          const uint32_t* count, uint64_t cap, int op, T s,
                          T* __restrict__ out, uint32_t* ovf_flag) {
  const uint64_t n = count ? (uint64_t)*count : cap;
  bool ovf = false;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    T x = col[sel ? sel[i] : i];
    if constexpr (IS_FP) out[i] = proj_fp<T>(x, s, op);
    else out[i] = proj_int<T>(x, s, op, ovf);
  }
  if constexpr (!IS_FP) {
    if (__ballot(ovf) && lane_id() == 0 && ovf_flag) atomicOr(ovf_flag, 1u);
  }
}

static uint32_t grid_for(uint64_t n) {
  uint64_t b = ceil_div(n ? n : 1, 256 * 4);
  uint64_t cap = (uint64_t)cu_count() * 16;
  return (uint32_t)(b < cap ? b : cap);
}

}  // namespace mtl

using namespace mtl;

extern "C" {

int32_t mtl_filter(const mtl_pred* preds, uint32_t n_preds, uint64_t n_rows,
                   const uint32_t* in_sel, const uint32_t* in_count,
                   uint64_t in_capacity, uint32_t* out_sel, uint32_t* out_count,
                   mtl_stream* s) {
  if (n_preds > MTL_MAX_PREDS)
    return set_error(MTL_UNSUPPORTED, "more than %d predicates", MTL_MAX_PREDS);
  if (!out_sel || !out_count)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "out_sel/out_count are NULL");
  if (n_preds && !preds) return set_error(MTL_ERR_INVALID_ARGUMENT, "preds is NULL");
  for (uint32_t p = 0; p < n_preds; ++p)
    if (!type_width(preds[p].type) || preds[p].type == MTL_HUGEINT ||
        preds[p].op < 0 || preds[p].op > 4 || !preds[p].column)
      return set_error(MTL_UNSUPPORTED, "predicate %u: unsupported type/op", p);
  uint64_t cap = in_sel ? in_capacity : n_rows;
  if (cap > 0xFFFFFFFFull)
    return set_error(MTL_UNSUPPORTED, "more than 2^32-1 rows");
  MTL_TRY(resolve_stream(s, &s));
  std::lock_guard<std::mutex> g(s->mu);
  s->arena.reset();
  uint32_t tiles = filter_tiles(cap);
  void *status, *ctr;
  MTL_TRY(s->arena.alloc((size_t)tiles * 8, &status));
  MTL_TRY(s->arena.alloc(256, &ctr));
  ZeroList zl;
  zl.add((uint32_t*)status, (uint64_t)tiles * 2);
  zl.add((uint32_t*)ctr, 64);
  MTL_TRY(launch_zero(zl, s->stream));
  FilterArgs a{};
  for (uint32_t p = 0; p < n_preds; ++p) a.preds[p] = preds[p];
  a.n_preds = n_preds;
  a.n_rows = cap;
  a.in_sel = in_sel;
  a.in_count = in_sel ? in_count : nullptr;
  a.out_sel = out_sel;
  a.out_count = out_count;
  a.status = (uint64_t*)status;
  a.tile_ctr = (uint32_t*)ctr;
  a.err = (uint32_t*)ctr + 1;
  ProfScope ps(s, "k_filter");
  return launch_filter(a, cap, s->stream);
}

int32_t mtl_compare(const mtl_pred* preds, uint32_t n_preds, uint64_t n_rows,
                    uint8_t* out_mask, mtl_stream* s) {
  if (n_preds > MTL_MAX_PREDS)
    return set_error(MTL_UNSUPPORTED, "more than %d predicates", MTL_MAX_PREDS);
  MTL_TRY(resolve_stream(s, &s));
  CompareArgs a{};
  for (uint32_t p = 0; p < n_preds; ++p) a.preds[p] = preds[p];
  a.n_preds = n_preds;
  a.n = n_rows;
  a.out = out_mask;
  if (n_rows == 0) return MTL_OK;
  hipLaunchKernelGGL(k_compare, dim3(grid_for(n_rows)), dim3(256), 0, s->stream, a);
  MTL_HIP(hipGetLastError());
  return MTL_OK;
}

int32_t mtl_gather(const void* column, uint32_t width, const uint32_t* sel,
                   const uint32_t* count, uint64_t capacity, void* out,
                   mtl_stream* s) {
  MTL_TRY(resolve_stream(s, &s));
  if (capacity == 0) return MTL_OK;
  dim3 g(grid_for(capacity)), b(256);
  switch (width) {
    case 1:
      hipLaunchKernelGGL(k_gather<uint8_t>, g, b, 0, s->stream,
                         (const uint8_t*)column, sel, count, capacity, (uint8_t*)out);
      break;
    case 4:
      hipLaunchKernelGGL(k_gather<uint32_t>, g, b, 0, s->stream,
                         (const uint32_t*)column, sel, count, capacity, (uint32_t*)out);
      break;
    case 8:
      hipLaunchKernelGGL(k_gather<uint64_t>, g, b, 0, s->stream,
                         (const uint64_t*)column, sel, count, capacity, (uint64_t*)out);
      break;
    default:
      return set_error(MTL_ERR_INVALID_ARGUMENT, "gather width %u", width);
  }
  MTL_HIP(hipGetLastError());
  return MTL_OK;
}

int32_t mtl_project(const void* column, int32_t type, const uint32_t* sel,
                    const uint32_t* count, uint64_t capacity, int32_t op,
                    int64_t scalar, void* out, uint32_t* overflow_flag,
                    mtl_stream* s) {
  if (op < 0 || op > 3) return set_error(MTL_ERR_INVALID_ARGUMENT, "project op %d", op);
  MTL_TRY(resolve_stream(s, &s));
  if (capacity == 0) return MTL_OK;
  dim3 g(grid_for(capacity)), b(256);
  switch (type) {
    case MTL_INTEGER:
      hipLaunchKernelGGL((k_project<int32_t, false>), g, b, 0, s->stream,
                         (const int32_t*)column, sel, count, capacity, op,
                         (int32_t)scalar, (int32_t*)out, overflow_flag);
      break;
    case MTL_DATE:
      // DATE +/- INTEGER days -> DATE (DuckDB has no DATE * INTEGER)
      if (op == MTL_PROJ_MUL || op == MTL_PROJ_SUBFROM)
        return set_error(MTL_UNSUPPORTED, "DATE supports only + and - of days");
      hipLaunchKernelGGL((k_project<int32_t, false>), g, b, 0, s->stream,
                         (const int32_t*)column, sel, count, capacity, op,
                         (int32_t)scalar, (int32_t*)out, overflow_flag);
      break;
    case MTL_BIGINT:
    case MTL_DECIMAL:
      hipLaunchKernelGGL((k_project<int64_t, false>), g, b, 0, s->stream,
                         (const int64_t*)column, sel, count, capacity, op, scalar,
                         (int64_t*)out, overflow_flag);
      break;
    case MTL_FLOAT:
      hipLaunchKernelGGL((k_project<float, true>), g, b, 0, s->stream,
                         (const float*)column, sel, count, capacity, op,
                         (float)__builtin_bit_cast(double, scalar), (float*)out,
                         overflow_flag);
      break;
    case MTL_DOUBLE:
      hipLaunchKernelGGL((k_project<double, true>), g, b, 0, s->stream,
                         (const double*)column, sel, count, capacity, op,
                         __builtin_bit_cast(double, scalar), (double*)out,
                         overflow_flag);
      break;
    default:
      return set_error(MTL_UNSUPPORTED, "projection on type %d", type);
  }
  MTL_HIP(hipGetLastError());
  return MTL_OK;
}

}  // extern "C"
