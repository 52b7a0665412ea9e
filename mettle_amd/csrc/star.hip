// Star join -> grouped SUM -> top-K, device resident end to end.
//
// The shape is TPC-H Q3 (and BASELINE config 3): a dimension D filtered to a
// key set, a parent P filtered and semi-joined to D, a fact F filtered and
// joined to P, SUM(expr) grouped by the P row, ORDER BY sum DESC + ties,
// LIMIT K.  Reference pipeline: execute_gpu_native
// (src/plan/executor/fast_path/mod.rs:58-141) with filter.rs:13-90,
// join_loop.rs:27-332, aggregate.rs:19-296, materialize.rs:19-80.
//
// MI355X plan (all HBM-bound integer work, no MFMA):
//   1. k_mark_dim : filter D, set bit (d_key - min) in a dimension bitmap
//                   (1 bit per key: 188 KB for 1.5M keys, L2 resident).
//   2. k_filter   : filter P + probe the D bitmap with p_fk, stable
//                   compaction of qualifying P rows, and set bit
//                   (p_key - min) in the parent bitmap.
//   3. k_rank     : exclusive popcount prefix over the parent bitmap words:
//                   rank(s) = dense accumulator index of key slot s.
//   4. k_map      : row_of[rank] = P row, clear the accumulators.
//   5. k_fact     : stream F's foreign key, probe the parent bitmap, and only
//                   for hits read the filter / expression columns; add the
//                   exact int64 (or f64) value to acc[rank] with an atomic.
//   6. k_topk     : per-block LDS top-K over (sum desc, ties, key asc), the
//                   last block merges and writes the K result rows.
// The bitmap+rank pair replaces the reference's 4-byte-per-key dense lookup
// (dense_lookup_build_i32_v, kernels.metal:839-852): 1.9 MB instead of 60 MB
// for 15M keys, so every probe of step 5 hits L2.
#include <algorithm>
#include <cstring>

#include "device.hpp"
#include "filter.hpp"
#include "mtl_internal.hpp"
#include "prof.hpp"
#include "topk.hpp"

namespace mtl {
using namespace dev;

struct DimArgs {
  mtl_pred preds[MTL_MAX_PREDS];
  uint32_t n_preds;
  uint64_t n;
  const void* key;
  int32_t key_type;
  int64_t min;
  uint64_t span;
  uint64_t* bm;
  uint32_t* err;
};

// err bits: 1 look-back spin cap, 2 duplicate parent key, 4 duplicate dim key,
// 8 key outside the stated [min, max], 16 expression overflow, 32 sum overflow
__global__ void k_mark_dim(DimArgs a) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    bool ok = true;
    for (uint32_t p = 0; p < a.n_preds && ok; ++p) ok = eval_pred_row(a.preds[p], i);
    if (!ok) continue;
    uint64_t s = key_off(load_key(a.key, a.key_type, i), a.min);
    if (s >= a.span) {
      atomicOr(a.err, 8u);
      continue;
    }
    unsigned long long bit = 1ull << (s & 63);
    unsigned long long old = atomicOr((unsigned long long*)&a.bm[s >> 6], bit);
    if (old & bit) atomicOr(a.err, 4u);
  }
}

// Exclusive popcount prefix over bitmap words (decoupled look-back).
constexpr int RBLOCK = 256;
constexpr int RPER = 8;
constexpr int RTILE = RBLOCK * RPER;

__global__ __launch_bounds__(RBLOCK) void k_rank(const uint64_t* __restrict__ bm,
                                                 uint64_t words,
                                                 uint32_t* __restrict__ rank,
                                                 uint64_t* status, uint32_t* ctr,
                                                 uint32_t* err) {
  __shared__ uint32_t s_tile, s_prefix;
  __shared__ uint32_t s_wave[RBLOCK / WAVE];
  const int tid = threadIdx.x, wave = tid / WAVE, lane = tid % WAVE;
  if (tid == 0) s_tile = atomicAdd(ctr, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t w0 = (uint64_t)tile * RTILE + (uint64_t)tid * RPER;
  uint32_t c[RPER];
  uint32_t sum = 0;
#pragma unroll
  for (int i = 0; i < RPER; ++i) {
    c[i] = (w0 + i < words) ? (uint32_t)__popcll(bm[w0 + i]) : 0u;
    sum += c[i];
  }
  uint32_t incl = wave_inclusive_u32(sum);
  if (lane == WAVE - 1) s_wave[wave] = incl;
  __syncthreads();
  uint32_t woff = 0, total = 0;
#pragma unroll
  for (int w = 0; w < RBLOCK / WAVE; ++w) {
    if (w < wave) woff += s_wave[w];
    total += s_wave[w];
  }
  if (wave == 0) {
    uint32_t ex = lookback_wave(status, tile, total, err);
    if (lane == 0) s_prefix = ex;
  }
  __syncthreads();
  uint32_t run = s_prefix + woff + incl - sum;
#pragma unroll
  for (int i = 0; i < RPER; ++i) {
    if (w0 + i < words) rank[w0 + i] = run;
    run += c[i];
  }
}

__device__ __forceinline__ uint32_t rank_of(const uint64_t* bm, const uint32_t* rank,
                                            uint64_t s) {
  uint64_t w = bm[s >> 6];
  uint64_t below = (s & 63) ? (w & ((1ull << (s & 63)) - 1)) : 0ull;
  return rank[s >> 6] + (uint32_t)__popcll(below);
}

struct MapArgs {
  const uint32_t* prow;
  const uint32_t* count;
  const void* key;
  int32_t key_type;
  int64_t min;
  uint64_t span;
  const uint64_t* bm;
  const uint32_t* rank;
  uint32_t* row_of;
  uint64_t* acc;
  uint32_t* cnt;
};

// Slots r < (distinct qualifying keys) <= n get their P row.  Accumulators
// are cleared by position i < n, which covers every r even when duplicate
// keys (flagged by the parent pass) make the distinct count smaller than n;
// slots without fact rows keep cnt == 0 and are never read by the top-K.
__global__ void k_map(MapArgs a) {
  const uint32_t n = *a.count;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    a.acc[i] = 0;
    a.cnt[i] = 0;
    uint32_t row = a.prow[i];
    uint64_t s = key_off(load_key(a.key, a.key_type, row), a.min);
    if (s >= a.span) continue;  // flagged by the parent pass
    uint32_t r = rank_of(a.bm, a.rank, s);
    a.row_of[r] = row;
  }
}

// ---------------------------------------------------------------------------
// Expression program (postfix), evaluated only for qualifying fact rows.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t expr_i64(const mtl_expr& e, uint64_t row,
                                            bool& ovf) {
  int64_t st[8];
  int sp = 0;
  for (uint32_t i = 0; i < e.n_ops; ++i) {
    int op = e.op[i];
    if (op == MTL_EXPR_COL) {
      int c = (int)e.arg[i];
      st[sp++] = load_key(e.col[c], e.col_type[c], row);
    } else if (op == MTL_EXPR_CONST) {
      st[sp++] = e.arg[i];
    } else {
      int64_t b = st[--sp], a = st[--sp], r;
      bool o;
      if (op == MTL_EXPR_ADD) o = __builtin_add_overflow(a, b, &r);
      else if (op == MTL_EXPR_SUB) o = __builtin_sub_overflow(a, b, &r);
      else o = __builtin_mul_overflow(a, b, &r);
      ovf |= o;
      st[sp++] = r;
    }
  }
  return sp ? st[sp - 1] : 0;
}

__device__ __forceinline__ double load_f64(const void* col, int32_t type, uint64_t row) {
  switch (type) {
    case MTL_FLOAT: return (double)((const float*)col)[row];
    case MTL_DOUBLE: return ((const double*)col)[row];
    default: return (double)load_key(col, type, row);
  }
}

__device__ __forceinline__ double expr_f64(const mtl_expr& e, uint64_t row) {
  double st[8];
  int sp = 0;
  for (uint32_t i = 0; i < e.n_ops; ++i) {
    int op = e.op[i];
    if (op == MTL_EXPR_COL) {
      int c = (int)e.arg[i];
      st[sp++] = load_f64(e.col[c], e.col_type[c], row);
    } else if (op == MTL_EXPR_CONST) {
      st[sp++] = __longlong_as_double(e.arg[i]);
    } else {
      double b = st[--sp], a = st[--sp];
      st[sp++] = op == MTL_EXPR_ADD ? a + b : (op == MTL_EXPR_SUB ? a - b : a * b);
    }
  }
  return sp ? st[sp - 1] : 0.0;
}

struct FactArgs {
  uint64_t n;
  const void* fk;
  int32_t fk_type;
  int64_t min;
  uint64_t span;
  const uint64_t* bm;
  const uint32_t* rank;
  mtl_pred preds[MTL_MAX_PREDS];
  uint32_t n_preds;
  mtl_expr expr;
  uint64_t* acc;  // int64 or double bits
  uint32_t* cnt;
  uint32_t* err;
  uint32_t fk_vec;  // 1: fk is 16-byte aligned
};

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void fact_row(const FactArgs& a, uint64_t row, int64_t key,
                                         bool& ovf, bool& sovf) {
  uint64_t s = key_off(key, a.min);
  if (s >= a.span) return;
  uint64_t w = a.bm[s >> 6];
  if (!((w >> (s & 63)) & 1ull)) return;
  for (uint32_t p = 0; p < a.n_preds; ++p)
    if (!eval_pred_row(a.preds[p], row)) return;
  uint64_t below = (s & 63) ? (w & ((1ull << (s & 63)) - 1)) : 0ull;
  uint32_t r = a.rank[s >> 6] + (uint32_t)__popcll(below);
  if (a.expr.is_double) {
    double v = expr_f64(a.expr, row);
    atomicAdd((double*)&a.acc[r], v);
  } else {
    int64_t v = expr_i64(a.expr, row, ovf);
    int64_t old = (int64_t)atomicAdd((unsigned long long*)&a.acc[r],
                                     (unsigned long long)v);
    int64_t t;
    sovf |= __builtin_add_overflow(old, v, &t);
  }
  atomicAdd(&a.cnt[r], 1u);
}

__global__ __launch_bounds__(256) void k_fact(FactArgs a) {
  bool ovf = false, sovf = false;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x * 4;
  for (uint64_t base = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
       base < a.n; base += stride) {
    int64_t key[4];
    if (a.fk_vec && base + 3 < a.n && a.fk_type != MTL_BIGINT &&
        a.fk_type != MTL_TIMESTAMP && a.fk_type != MTL_DECIMAL &&
        a.fk_type != MTL_DICT8) {
      u32x4 u = __builtin_nontemporal_load((const u32x4*)((const int32_t*)a.fk + base));
      key[0] = (int32_t)u.x;
      key[1] = (int32_t)u.y;
      key[2] = (int32_t)u.z;
      key[3] = (int32_t)u.w;
    } else if (a.fk_vec && base + 3 < a.n &&
               (a.fk_type == MTL_BIGINT || a.fk_type == MTL_TIMESTAMP ||
                a.fk_type == MTL_DECIMAL)) {
      u64x2 u0 = __builtin_nontemporal_load((const u64x2*)((const int64_t*)a.fk + base));
      u64x2 u1 =
          __builtin_nontemporal_load((const u64x2*)((const int64_t*)a.fk + base + 2));
      key[0] = (int64_t)u0.x;
      key[1] = (int64_t)u0.y;
      key[2] = (int64_t)u1.x;
      key[3] = (int64_t)u1.y;
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v)
        key[v] = base + v < a.n ? load_key(a.fk, a.fk_type, base + v) : 0;
    }
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (base + v < a.n) fact_row(a, base + v, key[v], ovf, sovf);
  }
  uint32_t e = (ovf ? 16u : 0u) | (sovf ? 32u : 0u);
  if (__ballot(e != 0) && e) atomicOr(a.err, e);
}

// ---------------------------------------------------------------------------
// Top-K over the accumulators.
// ---------------------------------------------------------------------------
struct TopkArgs {
  const uint32_t* m_count;  // number of accumulator slots (device)
  const uint64_t* acc;
  const uint32_t* cnt;
  const uint32_t* row_of;
  int32_t is_double;
  uint32_t k;
  uint32_t n_ties;
  const void* tie_col[MTL_STAR_MAX_TIES];
  int32_t tie_type[MTL_STAR_MAX_TIES];
  int32_t tie_desc[MTL_STAR_MAX_TIES];
  // per-block results
  uint64_t* ck[4];
  uint32_t* cid;
  uint32_t* cn;
  uint32_t* done;
  mtl_star_row* out;
  uint32_t* out_n;  // [0] = rows, [1] = error bits copied from err
  const uint32_t* err;
};

__device__ __forceinline__ void star_keys(const TopkArgs& a, uint32_t r, uint64_t k[4]) {
  uint64_t v = a.acc[r];
  k[0] = a.is_double ? ~enc_f64(__longlong_as_double((long long)v)) : ~enc_i64((int64_t)v);
  uint32_t row = a.row_of[r];
#pragma unroll
  for (int t = 0; t < MTL_STAR_MAX_TIES; ++t) {
    if ((uint32_t)t < a.n_ties) {
      uint64_t e = enc_value(a.tie_col[t], a.tie_type[t], row);
      k[t + 1] = a.tie_desc[t] ? ~e : e;
    } else {
      k[t + 1] = 0;
    }
  }
}

__global__ __launch_bounds__(TOPK_BLOCK) void k_star_topk(TopkArgs a) {
  __shared__ TopkLds s;
  __shared__ bool s_last;
  topk_init(s);
  const uint32_t m = *a.m_count;
  const uint32_t chunk = (m + gridDim.x - 1) / gridDim.x;
  const uint32_t start = blockIdx.x * chunk;
  const uint32_t end = min(m, start + chunk);
  for (uint32_t base = start; base < end; base += TOPK_ROUND) {
    for (int e = 0; e < 2; ++e) {
      uint32_t r = base + e * TOPK_BLOCK + threadIdx.x;
      if (r < end && a.cnt[r] > 0) {
        uint64_t v = a.acc[r];
        uint64_t k0 = a.is_double ? ~enc_f64(__longlong_as_double((long long)v))
                                  : ~enc_i64((int64_t)v);
        if (k0 <= s.thr) {
          uint64_t k[4];
          star_keys(a, r, k);
          topk_push(s, k[0], k[1], k[2], k[3], r);
        }
      }
    }
    topk_maybe_compact(s, a.k);
  }
  __syncthreads();
  uint32_t n = s.n;
  __syncthreads();
  topk_sort(s, n);
  topk_truncate(s, a.k);
  const uint32_t kb = s.n;
  for (uint32_t i = threadIdx.x; i < kb; i += blockDim.x) {
    uint64_t o = (uint64_t)blockIdx.x * a.k + i;
#pragma unroll
    for (int f = 0; f < 4; ++f) a.ck[f][o] = s.k[f][i];
    a.cid[o] = s.id[i];
  }
  if (threadIdx.x == 0) a.cn[blockIdx.x] = kb;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) s_last = atomicAdd(a.done, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  // merge: the same admission procedure over all block results
  topk_init(s);
  for (uint32_t b = 0; b < gridDim.x; ++b) {
    uint32_t nb = __hip_atomic_load(&a.cn[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (uint32_t base = 0; base < nb; base += TOPK_ROUND) {
      for (int e = 0; e < 2; ++e) {
        uint32_t i = base + e * TOPK_BLOCK + threadIdx.x;
        if (i < nb) {
          uint64_t o = (uint64_t)b * a.k + i;
          uint64_t k0 = a.ck[0][o];
          if (k0 <= s.thr) topk_push(s, k0, a.ck[1][o], a.ck[2][o], a.ck[3][o], a.cid[o]);
        }
      }
      topk_maybe_compact(s, a.k);
    }
  }
  __syncthreads();
  n = s.n;
  __syncthreads();
  topk_sort(s, n);
  topk_truncate(s, a.k);
  for (uint32_t i = threadIdx.x; i < s.n; i += blockDim.x) {
    uint32_t r = s.id[i];
    mtl_star_row row;
    row.p_row = a.row_of[r];
    row.count = a.cnt[r];
    uint64_t v = a.acc[r];
    row.sum_i64 = a.is_double ? 0 : (int64_t)v;
    row.sum_f64 = a.is_double ? __longlong_as_double((long long)v) : (double)(int64_t)v;
    a.out[i] = row;
  }
  if (threadIdx.x == 0) {
    a.out_n[0] = s.n;
    a.out_n[1] = __hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
static bool int_key_type(int32_t t) {
  return t == MTL_INTEGER || t == MTL_BIGINT || t == MTL_DATE ||
         t == MTL_TIMESTAMP || t == MTL_DECIMAL || t == MTL_DICT32 || t == MTL_DICT8;
}

static int32_t check_preds(const mtl_pred* p, uint32_t n, const char* what) {
  if (n > MTL_MAX_PREDS) return set_error(MTL_UNSUPPORTED, "%s: too many predicates", what);
  for (uint32_t i = 0; i < n; ++i)
    if (!p[i].column || !type_width(p[i].type) || p[i].type == MTL_HUGEINT ||
        p[i].op < 0 || p[i].op > 4)
      return set_error(MTL_UNSUPPORTED, "%s: predicate %u unsupported", what, i);
  return MTL_OK;
}

static int32_t check_expr(const mtl_expr& e) {
  if (e.n_ops == 0 || e.n_ops > MTL_EXPR_MAX_OPS || e.n_cols > MTL_EXPR_MAX_COLS)
    return set_error(MTL_UNSUPPORTED, "aggregate expression has %u ops", e.n_ops);
  int depth = 0;
  for (uint32_t i = 0; i < e.n_ops; ++i) {
    switch (e.op[i]) {
      case MTL_EXPR_COL:
        if (e.arg[i] < 0 || (uint64_t)e.arg[i] >= e.n_cols || !e.col[e.arg[i]])
          return set_error(MTL_ERR_INVALID_ARGUMENT, "expression column %lld",
                           (long long)e.arg[i]);
        if (!e.is_double && !int_key_type(e.col_type[e.arg[i]]))
          return set_error(MTL_UNSUPPORTED, "integer expression over a float column");
        ++depth;
        break;
      case MTL_EXPR_CONST:
        ++depth;
        break;
      case MTL_EXPR_ADD:
      case MTL_EXPR_SUB:
      case MTL_EXPR_MUL:
        if (depth < 2) return set_error(MTL_ERR_INVALID_ARGUMENT, "expression stack underflow");
        --depth;
        break;
      default:
        return set_error(MTL_ERR_INVALID_ARGUMENT, "expression op %d", e.op[i]);
    }
    if (depth > 8) return set_error(MTL_UNSUPPORTED, "expression deeper than 8");
  }
  if (depth != 1) return set_error(MTL_ERR_INVALID_ARGUMENT, "expression leaves %d values", depth);
  return MTL_OK;
}

struct StarPlan {
  uint64_t dspan = 0, pspan = 0, pwords = 0, dwords = 0;
  uint32_t topk_grid = 0;
};

static int32_t star_validate(const mtl_star_desc* d, StarPlan* sp) {
  if (!d) return set_error(MTL_ERR_INVALID_ARGUMENT, "desc is NULL");
  if (d->k == 0 || d->k > (uint32_t)TOPK_MAX_K)
    return set_error(MTL_UNSUPPORTED, "top-K with K=%u (supported 1..%d)", d->k, TOPK_MAX_K);
  if (d->n_ties > MTL_STAR_MAX_TIES)
    return set_error(MTL_UNSUPPORTED, "more than %d tie-break columns", MTL_STAR_MAX_TIES);
  for (uint32_t t = 0; t < d->n_ties; ++t)
    if (!d->tie_col[t] || !type_width(d->tie_type[t]) || d->tie_type[t] == MTL_HUGEINT)
      return set_error(MTL_UNSUPPORTED, "tie-break column %u unsupported", t);
  if (d->p_rows > 0xFFFFFFFFull || d->f_rows > 0xFFFFFFFFull || d->d_rows > 0xFFFFFFFFull)
    return set_error(MTL_UNSUPPORTED, "more than 2^32-1 rows");
  if (!d->p_key || !int_key_type(d->p_key_type) || !d->f_fk || !int_key_type(d->f_fk_type))
    return set_error(MTL_UNSUPPORTED, "join keys must be integer columns");
  if (d->d_rows && (!d->d_key || !int_key_type(d->d_key_type) || !d->p_fk ||
                    !int_key_type(d->p_fk_type)))
    return set_error(MTL_UNSUPPORTED, "dimension keys must be integer columns");
  MTL_TRY(check_preds(d->d_preds, d->d_rows ? d->d_n_preds : 0, "dimension"));
  MTL_TRY(check_preds(d->p_preds, d->p_n_preds, "parent"));
  MTL_TRY(check_preds(d->f_preds, d->f_n_preds, "fact"));
  MTL_TRY(check_expr(d->agg));
  if (d->p_key_max < d->p_key_min) return set_error(MTL_ERR_INVALID_ARGUMENT, "parent key range");
  sp->pspan = key_off(d->p_key_max, d->p_key_min) + 1;
  if (sp->pspan == 0 || sp->pspan > (1ull << 36))
    return set_error(MTL_UNSUPPORTED, "parent key span %llu too wide for a bitmap",
                     (unsigned long long)sp->pspan);
  sp->pwords = ceil_div(sp->pspan, 64);
  if (d->d_rows) {
    if (d->d_key_max < d->d_key_min) return set_error(MTL_ERR_INVALID_ARGUMENT, "dim key range");
    sp->dspan = key_off(d->d_key_max, d->d_key_min) + 1;
    if (sp->dspan == 0 || sp->dspan > (1ull << 36))
      return set_error(MTL_UNSUPPORTED, "dimension key span too wide for a bitmap");
    sp->dwords = ceil_div(sp->dspan, 64);
  }
  uint64_t g = ceil_div(d->p_rows ? d->p_rows : 1, 8192);
  sp->topk_grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, (uint64_t)cu_count()));
  return MTL_OK;
}

// Enqueue the whole pipeline on s.  out (device, k rows) and out_n (device,
// 2 x uint32: rows, error bits) receive the result.
static int32_t star_enqueue(const mtl_star_desc* d, const StarPlan& sp, mtl_stream* s,
                            mtl_star_row* out, uint32_t* out_n) {
  hipStream_t st = s->stream;
  Arena& ar = s->arena;
  const uint64_t prows = d->p_rows;
  void *dim_bm = nullptr, *par_bm, *par_rank, *prow, *row_of, *acc, *cnt, *misc;
  void *st_p, *st_r, *ck[4], *cid, *cn;
  if (d->d_rows) MTL_TRY(ar.alloc(sp.dwords * 8, &dim_bm));
  MTL_TRY(ar.alloc(sp.pwords * 8, &par_bm));
  MTL_TRY(ar.alloc(sp.pwords * 4, &par_rank));
  MTL_TRY(ar.alloc((prows ? prows : 1) * 4, &prow));
  MTL_TRY(ar.alloc((prows ? prows : 1) * 4, &row_of));
  MTL_TRY(ar.alloc((prows ? prows : 1) * 8, &acc));
  MTL_TRY(ar.alloc((prows ? prows : 1) * 4, &cnt));
  MTL_TRY(ar.alloc(256, &misc));  // [0]=count_p [1]=tile ctr P [2]=tile ctr rank
                                  // [3]=err [4]=done
  uint32_t ptiles = filter_tiles(prows);
  uint32_t rtiles = (uint32_t)ceil_div(sp.pwords, RTILE);
  MTL_TRY(ar.alloc((uint64_t)ptiles * 8, &st_p));
  MTL_TRY(ar.alloc((uint64_t)std::max<uint32_t>(rtiles, 1) * 8, &st_r));
  for (int f = 0; f < 4; ++f) MTL_TRY(ar.alloc((uint64_t)sp.topk_grid * d->k * 8, &ck[f]));
  MTL_TRY(ar.alloc((uint64_t)sp.topk_grid * d->k * 4, &cid));
  MTL_TRY(ar.alloc((uint64_t)sp.topk_grid * 4, &cn));
  uint32_t* mu = (uint32_t*)misc;

  ZeroList zl;
  if (dim_bm) zl.add((uint32_t*)dim_bm, sp.dwords * 2);
  zl.add((uint32_t*)par_bm, sp.pwords * 2);
  zl.add((uint32_t*)misc, 64);
  zl.add((uint32_t*)st_p, (uint64_t)ptiles * 2);
  zl.add((uint32_t*)st_r, (uint64_t)std::max<uint32_t>(rtiles, 1) * 2);
  MTL_TRY(launch_zero(zl, st));

  // 1. dimension bitmap
  if (d->d_rows) {
    DimArgs da{};
    for (uint32_t p = 0; p < d->d_n_preds; ++p) da.preds[p] = d->d_preds[p];
    da.n_preds = d->d_n_preds;
    da.n = d->d_rows;
    da.key = d->d_key;
    da.key_type = d->d_key_type;
    da.min = d->d_key_min;
    da.span = sp.dspan;
    da.bm = (uint64_t*)dim_bm;
    da.err = mu + 3;
    uint32_t g = (uint32_t)std::min<uint64_t>(ceil_div(d->d_rows, 256 * 4), 4096);
    ProfScope ps(s, "k_mark_dim");
    hipLaunchKernelGGL(k_mark_dim, dim3(std::max<uint32_t>(g, 1)), dim3(256), 0, st, da);
    MTL_HIP(hipGetLastError());
  }
  // 2. parent filter + semi-join probe + compaction + parent bitmap
  {
    FilterArgs fa{};
    for (uint32_t p = 0; p < d->p_n_preds; ++p) fa.preds[p] = d->p_preds[p];
    fa.n_preds = d->p_n_preds;
    fa.n_rows = prows;
    fa.out_sel = (uint32_t*)prow;
    fa.out_count = mu + 0;
    fa.status = (uint64_t*)st_p;
    fa.tile_ctr = mu + 1;
    fa.err = mu + 3;
    if (d->d_rows) {
      fa.probe_bm = (const uint64_t*)dim_bm;
      fa.probe_key = d->p_fk;
      fa.probe_key_type = d->p_fk_type;
      fa.probe_min = d->d_key_min;
      fa.probe_span = sp.dspan;
    }
    fa.mark_bm = (uint64_t*)par_bm;
    fa.mark_key = d->p_key;
    fa.mark_key_type = d->p_key_type;
    fa.mark_min = d->p_key_min;
    fa.mark_span = sp.pspan;
    ProfScope ps(s, "k_filter_parent");
    MTL_TRY(launch_filter(fa, prows, st));
  }
  // 3. rank
  {
    ProfScope ps(s, "k_rank");
    hipLaunchKernelGGL(k_rank, dim3(std::max<uint32_t>(rtiles, 1)), dim3(RBLOCK), 0, st,
                       (const uint64_t*)par_bm, sp.pwords, (uint32_t*)par_rank,
                       (uint64_t*)st_r, mu + 2, mu + 3);
  }
  MTL_HIP(hipGetLastError());
  // 4. map accumulator slots to P rows, clear accumulators
  {
    MapArgs ma{};
    ma.prow = (const uint32_t*)prow;
    ma.count = mu + 0;
    ma.key = d->p_key;
    ma.key_type = d->p_key_type;
    ma.min = d->p_key_min;
    ma.span = sp.pspan;
    ma.bm = (const uint64_t*)par_bm;
    ma.rank = (const uint32_t*)par_rank;
    ma.row_of = (uint32_t*)row_of;
    ma.acc = (uint64_t*)acc;
    ma.cnt = (uint32_t*)cnt;
    uint32_t g = (uint32_t)std::min<uint64_t>(ceil_div(prows ? prows : 1, 256 * 4), 2048);
    ProfScope ps(s, "k_map");
    hipLaunchKernelGGL(k_map, dim3(std::max<uint32_t>(g, 1)), dim3(256), 0, st, ma);
    MTL_HIP(hipGetLastError());
  }
  // 5. fact pass
  if (d->f_rows) {
    FactArgs fa{};
    fa.n = d->f_rows;
    fa.fk = d->f_fk;
    fa.fk_type = d->f_fk_type;
    fa.min = d->p_key_min;
    fa.span = sp.pspan;
    fa.bm = (const uint64_t*)par_bm;
    fa.rank = (const uint32_t*)par_rank;
    for (uint32_t p = 0; p < d->f_n_preds; ++p) fa.preds[p] = d->f_preds[p];
    fa.n_preds = d->f_n_preds;
    fa.expr = d->agg;
    fa.acc = (uint64_t*)acc;
    fa.cnt = (uint32_t*)cnt;
    fa.err = mu + 3;
    fa.fk_vec = ((uintptr_t)d->f_fk % 16) == 0;
    uint64_t g = ceil_div(d->f_rows, 256 * 4 * 4);
    uint64_t gmax = (uint64_t)cu_count() * 32;
    ProfScope ps(s, "k_fact");
    hipLaunchKernelGGL(k_fact, dim3((uint32_t)std::max<uint64_t>(1, std::min(g, gmax))),
                       dim3(256), 0, st, fa);
    MTL_HIP(hipGetLastError());
  }
  // 6. top-K + output rows
  {
    TopkArgs ta{};
    ta.m_count = mu + 0;
    ta.acc = (const uint64_t*)acc;
    ta.cnt = (const uint32_t*)cnt;
    ta.row_of = (const uint32_t*)row_of;
    ta.is_double = d->agg.is_double;
    ta.k = d->k;
    ta.n_ties = d->n_ties;
    for (uint32_t t = 0; t < d->n_ties; ++t) {
      ta.tie_col[t] = d->tie_col[t];
      ta.tie_type[t] = d->tie_type[t];
      ta.tie_desc[t] = d->tie_desc[t];
    }
    for (int f = 0; f < 4; ++f) ta.ck[f] = (uint64_t*)ck[f];
    ta.cid = (uint32_t*)cid;
    ta.cn = (uint32_t*)cn;
    ta.done = mu + 4;
    ta.out = out;
    ta.out_n = out_n;
    ta.err = mu + 3;
    ProfScope ps(s, "k_star_topk");
    hipLaunchKernelGGL(k_star_topk, dim3(sp.topk_grid), dim3(TOPK_BLOCK), 0, st, ta);
    MTL_HIP(hipGetLastError());
  }
  return MTL_OK;
}

static int32_t star_error(uint32_t err) {
  if (err & 1u) return set_error(MTL_ERR_PROCESSING, "processing error: look-back wait exceeded");
  if (err & 8u)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "a key lies outside its stated [min, max]");
  if (err & (2u | 4u))
    return set_error(MTL_UNSUPPORTED, "join build key is not unique (fallback)");
  if (err & 16u) return set_error(MTL_ERR_OVERFLOW, "Overflow in aggregate expression");
  if (err & 32u) return set_error(MTL_UNSUPPORTED, "int64 SUM overflow (needs HUGEINT)");
  return MTL_OK;
}

int32_t star_check_error(uint32_t err) { return star_error(err); }

}  // namespace mtl

using namespace mtl;

extern "C" {

int32_t mtl_star_topk_async(const mtl_star_desc* desc, void* out_dev, uint32_t* out_n_dev,
                            mtl_stream* s) {
  StarPlan sp;
  MTL_TRY(star_validate(desc, &sp));
  if (!out_dev || !out_n_dev) return set_error(MTL_ERR_INVALID_ARGUMENT, "outputs are NULL");
  MTL_TRY(resolve_stream(s, &s));
  std::lock_guard<std::mutex> g(s->mu);
  s->arena.reset();
  return star_enqueue(desc, sp, s, (mtl_star_row*)out_dev, out_n_dev);
}

int32_t mtl_star_topk(const mtl_star_desc* desc, mtl_star_row* out_rows, uint32_t* out_n,
                      mtl_stream* s) {
  StarPlan sp;
  MTL_TRY(star_validate(desc, &sp));
  if (!out_rows || !out_n) return set_error(MTL_ERR_INVALID_ARGUMENT, "outputs are NULL");
  MTL_TRY(resolve_stream(s, &s));
  std::lock_guard<std::mutex> g(s->mu);
  s->arena.reset();
  void *dout, *dn, *host;
  size_t rows_bytes = (size_t)desc->k * sizeof(mtl_star_row);
  MTL_TRY(s->arena.alloc(rows_bytes, &dout));
  MTL_TRY(s->arena.alloc(16, &dn));
  MTL_TRY(star_enqueue(desc, sp, s, (mtl_star_row*)dout, (uint32_t*)dn));
  MTL_TRY(stream_pinned(s, rows_bytes + 16, &host));
  MTL_HIP(hipMemcpyAsync(host, dn, 8, hipMemcpyDeviceToHost, s->stream));
  MTL_HIP(hipMemcpyAsync((char*)host + 16, dout, rows_bytes, hipMemcpyDeviceToHost, s->stream));
  MTL_HIP(hipStreamSynchronize(s->stream));
  uint32_t n = ((uint32_t*)host)[0], err = ((uint32_t*)host)[1];
  MTL_TRY(star_error(err));
  memcpy(out_rows, (char*)host + 16, (size_t)n * sizeof(mtl_star_row));
  *out_n = n;
  return MTL_OK;
}

}  // extern "C"

extern "C" int32_t mtl_star_status(uint32_t error_bits) {
  return mtl::star_check_error(error_bits);
}
