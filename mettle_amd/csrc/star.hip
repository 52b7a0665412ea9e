// Star join -> grouped SUM -> top-K, device resident end to end.
//
// The shape is TPC-H Q3 (customers / orders / lines): a dimension D filtered to a
// key set, a parent P filtered and semi-joined to D, a fact F filtered and
// joined to P, SUM(expr) grouped by the P row, ORDER BY sum DESC + ties,
// LIMIT K.  Reference pipeline: execute_gpu_native
// (src/plan/executor/fast_path/mod.rs:58-141) with filter.rs:13-90,
// join_loop.rs:27-332, aggregate.rs:19-296, materialize.rs:19-80.
//
// MI355X plan (all HBM-bound integer work, no MFMA):
//   1. k_filter   : (mark-only) filter D, set bit (d_key - min) in a bitmap
//                   (1 bit per key: 188 KB for 1.5M keys, L2 resident).
//   2. k_filter   : filter P + probe the D bitmap with p_fk, write the pass
//                   bits (1 per P row) and set bit (p_key - min) in the
//                   parent bitmap.  No compaction: nothing downstream needs
//                   a selection vector of P.
//   3. k_rank     : exclusive popcount prefix over the parent bitmap words:
//                   rank(s) = dense accumulator index of key slot s.
//   4. k_map      : walk the pass bits: row_of[rank] = P row; clear the
//                   accumulators.
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
#include <type_traits>

#include "device.hpp"
#include "filter.hpp"
#include "mtl_internal.hpp"
#include "prof.hpp"
#include "rowtile.hpp"
#include "star.hpp"
#include "topk.hpp"

namespace mtl {
using namespace dev;

// err bits: 1 look-back spin cap, 2 duplicate parent key, 4 duplicate dim key,
// 8 key outside the stated [min, max], 16 expression overflow, 32 sum overflow

// Exclusive popcount prefix over bitmap words (decoupled look-back).
constexpr int RBLOCK = 256;
constexpr int RPER = 8;
constexpr int RTILE = RBLOCK * RPER;

// Also writes the parent bitmap's total popcount (*par_total, last tile) and
// accumulates three popcounts for the duplicate-key checks of k_map, with
// one atomic per block (grid-stride): pops[0] += |pass bits of P|,
// pops[1] += |dimension bitmap|, pops[2] += |pass bits of D|.
struct PopArgs {
  const uint64_t* pbits;
  uint64_t pbits_words;
  const uint64_t* dim_bm;
  uint64_t dim_words;
  const uint64_t* dbits;
  uint64_t dbits_words;
  uint32_t* pops;
};

// This thread's part of a grid-stride popcount, eight loads in flight.
__device__ __forceinline__ uint32_t thread_popcount(const uint64_t* p, uint64_t words,
                                                    uint32_t bid, uint32_t nb) {
  uint32_t c = 0;
  const uint64_t stride = (uint64_t)nb * RBLOCK;
  for (uint64_t i0 = (uint64_t)bid * RBLOCK + threadIdx.x; i0 < words; i0 += stride * 8) {
    uint64_t x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint64_t i = i0 + u * stride;
      x[u] = i < words ? p[i] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) c += (uint32_t)__popcll(x[u]);
  }
  return c;
}

__device__ __forceinline__ uint32_t block_sum(uint32_t c, uint32_t* s_red) {
  c = wave_sum_u32(c);
  __syncthreads();
  if (threadIdx.x % WAVE == 0) s_red[threadIdx.x / WAVE] = c;
  __syncthreads();
  uint32_t t = 0;
  for (int w = 0; w < RBLOCK / WAVE; ++w) t += s_red[w];
  return t;
}

// Blocks past the rank tiles (tickets >= rtiles) only add popcounts: the
// grid covers the larger of the parent bitmap and the dimension bitmaps, so
// a small parent table with a large dimension table is not left to a few
// blocks.
__device__ __forceinline__ void rank_body(const uint64_t* __restrict__ bm, uint64_t words,
                                          uint32_t rtiles, uint32_t* __restrict__ rank,
                                          uint64_t* status, uint32_t* ctr, uint32_t* err,
                                          uint32_t* par_total, const PopArgs& pa, uint32_t bid,
                                          uint32_t nb) {
  __shared__ uint32_t s_tile, s_prefix;
  __shared__ uint32_t s_wave[RBLOCK / WAVE];
  __shared__ uint32_t s_red[RBLOCK / WAVE];
  const int tid = threadIdx.x, wave = tid / WAVE, lane = tid % WAVE;
  if (tid == 0) s_tile = atomicAdd(ctr, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const bool ranks = tile < rtiles;  // block-uniform
  const uint64_t w0 = (uint64_t)tile * RTILE + (uint64_t)tid * RPER;
  uint64_t x[RPER];
#pragma unroll
  for (int i = 0; i < RPER; ++i) x[i] = (ranks && w0 + i < words) ? bm[w0 + i] : 0ull;
  // the duplicate-key popcounts are independent of the prefix: their loads
  // go out with the bitmap words and finish while predecessors publish
  const uint32_t p0 = thread_popcount(pa.pbits, pa.pbits_words, bid, nb);
  const uint32_t p1 = pa.dim_bm ? thread_popcount(pa.dim_bm, pa.dim_words, bid, nb) : 0u;
  const uint32_t p2 = pa.dbits ? thread_popcount(pa.dbits, pa.dbits_words, bid, nb) : 0u;
  if (ranks) {
    uint32_t c[RPER];
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < RPER; ++i) {
      c[i] = (uint32_t)__popcll(x[i]);
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
    if (tile == rtiles - 1 && tid == 0) *par_total = s_prefix + total;
  }
  const uint32_t c0 = block_sum(p0, s_red);
  const uint32_t c1 = block_sum(p1, s_red);
  const uint32_t c2 = block_sum(p2, s_red);
  if (tid == 0) {
    if (c0) atomicAdd(&pa.pops[0], c0);
    if (c1) atomicAdd(&pa.pops[1], c1);
    if (c2) atomicAdd(&pa.pops[2], c2);
  }
}

__global__ __launch_bounds__(RBLOCK) void k_rank(const uint64_t* __restrict__ bm,
                                                 uint64_t words, uint32_t rtiles,
                                                 uint32_t* __restrict__ rank,
                                                 uint64_t* status, uint32_t* ctr,
                                                 uint32_t* err, uint32_t* par_total,
                                                 PopArgs pa) {
  rank_body(bm, words, rtiles, rank, status, ctr, err, par_total, pa, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ uint32_t rank_of(const uint64_t* bm, const uint32_t* rank,
                                            uint64_t s) {
  uint64_t w = bm[s >> 6];
  uint64_t below = (s & 63) ? (w & ((1ull << (s & 63)) - 1)) : 0ull;
  return rank[s >> 6] + (uint32_t)__popcll(below);
}

struct MapArgs {
  const uint64_t* sel_bm;  // parent pass bits, 1 per P row
  uint64_t n_rows;
  const void* key;
  int32_t key_type;
  int64_t min;
  uint64_t span;
  const uint64_t* bm;
  const uint32_t* rank;
  uint32_t* row_of;
  uint64_t* acc;
  uint32_t* cnt;
  const uint32_t* par_total;  // parent bitmap popcount = accumulator slots
  const uint32_t* pops;       // [0] P pass bits, [1] D bitmap, [2] D pass bits
  uint32_t has_dim;
  uint32_t* err;
  uint32_t grid;  // blocks of this map (shared multi-query launches)
};

// Every qualifying P row (a set pass bit) writes row_of[rank(key)].  Slots
// r < par_total are exactly the set bits of the parent bitmap, each set by
// at least one qualifying row, so every slot gets its row; the accumulators
// are cleared by index.  A repeated key (par_total < count) is flagged.
// Thread t owns 16 consecutive rows (one 16-bit piece of the pass bits).
__device__ __forceinline__ void map_body(const MapArgs& a, uint32_t bid, uint32_t nb) {
  const uint32_t M = *a.par_total;
  if (bid == 0 && threadIdx.x == 0) {
    if (M != a.pops[0]) atomicOr(a.err, 2u);                     // repeated parent key
    if (a.has_dim && a.pops[1] != a.pops[2]) atomicOr(a.err, 4u);  // repeated dim key
  }
  const uint64_t nthreads = (uint64_t)nb * blockDim.x;
  const uint64_t t0 = (uint64_t)bid * blockDim.x + threadIdx.x;
  for (uint64_t i = t0; i < M; i += nthreads) {
    a.acc[i] = 0;
    a.cnt[i] = 0;
  }
  // 16 consecutive rows per thread (one 16-bit piece of the pass bits);
  // measured faster than issuing the key and rank loads of all qualifying
  // rows together, which costs registers (about 10% of rows qualify)
  for (uint64_t g = t0; g * 16 < a.n_rows; g += nthreads) {
    const uint64_t r0 = g * 16;
    uint32_t m = (uint32_t)(a.sel_bm[r0 / 64] >> (r0 % 64)) & 0xFFFFu;
    while (m) {
      int i = __ffs(m) - 1;
      m &= m - 1;
      uint64_t row = r0 + (uint64_t)i;
      uint64_t s = key_off(load_key(a.key, a.key_type, row), a.min);
      if (s >= a.span) continue;  // flagged by the parent pass
      a.row_of[rank_of(a.bm, a.rank, s)] = (uint32_t)row;
    }
  }
}

__global__ void k_map(MapArgs a) { map_body(a, blockIdx.x, gridDim.x); }

// 4-byte keys aligned for vector loads: a thread's 16 keys are one 64-byte
// line, loaded whole (4 vector loads) only when one of its rows qualifies;
// the slot lookups of its qualifying rows then go out four at a time (rows
// picked from the key registers with a select chain, unused lookups read
// word 0), so a thread waits about 1 + ceil(q / 4) round trips for q
// qualifying rows instead of 2 q.
__device__ __forceinline__ void map4_body(const MapArgs& a, uint32_t bid, uint32_t nb) {
  const uint32_t M = *a.par_total;
  if (bid == 0 && threadIdx.x == 0) {
    if (M != a.pops[0]) atomicOr(a.err, 2u);                     // repeated parent key
    if (a.has_dim && a.pops[1] != a.pops[2]) atomicOr(a.err, 4u);  // repeated dim key
  }
  const uint64_t nthreads = (uint64_t)nb * blockDim.x;
  const uint64_t t0 = (uint64_t)bid * blockDim.x + threadIdx.x;
  for (uint64_t i = t0; i < M; i += nthreads) {
    a.acc[i] = 0;
    a.cnt[i] = 0;
  }
  const uint32_t* key = (const uint32_t*)a.key;
  for (uint64_t g = t0; g * 16 < a.n_rows; g += nthreads) {
    const uint64_t r0 = g * 16;
    uint32_t m = (uint32_t)(a.sel_bm[r0 / 64] >> (r0 % 64)) & 0xFFFFu;
    if (!m) continue;
    uint32_t kv[16];
    if (r0 + 16 <= a.n_rows) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint4 u = *(const uint4*)(key + r0 + 4 * q);
        kv[4 * q] = u.x; kv[4 * q + 1] = u.y; kv[4 * q + 2] = u.z; kv[4 * q + 3] = u.w;
      }
    } else {
#pragma unroll
      for (int i = 0; i < 16; ++i) kv[i] = r0 + i < a.n_rows ? key[r0 + i] : 0u;
    }
    while (m) {
      uint32_t sl[4], rw[4], rk[4];
      uint64_t bw[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        sl[q] = 0xFFFFFFFFu;
        rw[q] = 0;
        if (m) {
          const int i = __ffs(m) - 1;
          m &= m - 1;
          uint32_t kk = kv[0];
#pragma unroll
          for (int t = 1; t < 16; ++t) kk = i == t ? kv[t] : kk;
          const uint64_t o = key_off((int64_t)(int32_t)kk, a.min);
          if (o < a.span) sl[q] = (uint32_t)o;  // else flagged by the parent pass
          rw[q] = (uint32_t)(r0 + (uint64_t)i);
        }
        const uint32_t wi = sl[q] != 0xFFFFFFFFu ? sl[q] >> 6 : 0u;
        bw[q] = a.bm[wi];
        rk[q] = a.rank[wi];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (sl[q] == 0xFFFFFFFFu) continue;
        const uint32_t o = sl[q];
        const uint64_t below = (o & 63) ? (bw[q] & ((1ull << (o & 63)) - 1)) : 0ull;
        a.row_of[rk[q] + (uint32_t)__popcll(below)] = rw[q];
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_map4(MapArgs a) { map4_body(a, blockIdx.x, gridDim.x); }

// The maps of a batch of queries: blockIdx.y is the query.
template <bool K4>
__global__ __launch_bounds__(256) void k_mq_map(const MapArgs* qs) {
  const MapArgs& a = qs[blockIdx.y];
  if (blockIdx.x >= a.grid) return;
  if constexpr (K4) map4_body(a, blockIdx.x, a.grid);
  else map_body(a, blockIdx.x, a.grid);
}

// ---------------------------------------------------------------------------
// Expression program (postfix), evaluated only for qualifying fact rows.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double load_f64(const void* col, int32_t type, uint64_t row) {
  switch (type) {
    case MTL_FLOAT: return (double)((const float*)col)[row];
    case MTL_DOUBLE: return ((const double*)col)[row];
    default: return (double)load_key(col, type, row);
  }
}

struct FactArgs {
  uint64_t n;
  const void* fk;
  int32_t fk_type;
  int64_t min;
  uint64_t span;
  const uint64_t* bm;
  const uint32_t* rank;
  PredC preds[MTL_MAX_PREDS];
  uint32_t n_preds;
  mtl_expr expr;
  uint32_t expr_w8;  // every expression column is 8 bytes wide (unused ones name column 0)
  uint64_t* acc;  // int64 or double bits
  uint32_t* cnt;
  uint32_t* err;
  uint32_t fk_vec;  // 1: fk is 16-byte aligned
  uint32_t pred0_vec;  // 1: the first predicate's column is 16-byte aligned
  uint32_t* hit_row;    // hit buffer: 1024-entry region per wave of each tile
  uint32_t* hit_slot;   // key slot of each hit
  uint32_t* hit_count;  // hits per wave region
  uint64_t n_regions;   // wave regions (tiles x 4)
  unsigned long long* tmax;  // per k_fact_acc block: max |value| it added
};

// Expression over preloaded column values (one row).  The operand stack
// lives in eight registers that shift on push and pop (the host checks
// depth <= 8), and column values are picked with a static select, so no
// dynamically indexed array spills to scratch memory.
template <typename V>
__device__ __forceinline__ V pick_col(const V (&cv)[MTL_EXPR_MAX_COLS], int64_t c) {
  V v = cv[0];
#pragma unroll
  for (int q = 1; q < MTL_EXPR_MAX_COLS; ++q)
    if (c == q) v = cv[q];
  return v;
}

template <typename V>
struct ShiftStack {
  V s[8];
  __device__ __forceinline__ void push(V v) {
#pragma unroll
    for (int i = 7; i > 0; --i) s[i] = s[i - 1];
    s[0] = v;
  }
  // replace the top two entries (s1 op s0) by r
  __device__ __forceinline__ void fold(V r) {
    s[0] = r;
#pragma unroll
    for (int i = 1; i < 7; ++i) s[i] = s[i + 1];
  }
};

__device__ __forceinline__ int64_t expr_i64_vals(const mtl_expr& e,
                                                 const int64_t (&cv)[MTL_EXPR_MAX_COLS],
                                                 bool& ovf) {
  ShiftStack<int64_t> st;
  for (uint32_t i = 0; i < e.n_ops; ++i) {
    const int op = e.op[i];
    if (op == MTL_EXPR_COL) {
      st.push(pick_col(cv, e.arg[i]));
    } else if (op == MTL_EXPR_CONST) {
      st.push(e.arg[i]);
    } else {
      const int64_t y = st.s[0], x = st.s[1];
      int64_t r;
      bool o;
      if (op == MTL_EXPR_ADD) o = __builtin_add_overflow(x, y, &r);
      else if (op == MTL_EXPR_SUB) o = __builtin_sub_overflow(x, y, &r);
      else o = __builtin_mul_overflow(x, y, &r);
      ovf |= o;
      st.fold(r);
    }
  }
  return st.s[0];
}

__device__ __forceinline__ double expr_f64_vals(const mtl_expr& e,
                                                const double (&cv)[MTL_EXPR_MAX_COLS]) {
  ShiftStack<double> st;
  for (uint32_t i = 0; i < e.n_ops; ++i) {
    const int op = e.op[i];
    if (op == MTL_EXPR_COL) {
      st.push(pick_col(cv, e.arg[i]));
    } else if (op == MTL_EXPR_CONST) {
      st.push(__longlong_as_double(e.arg[i]));
    } else {
      const double y = st.s[0], x = st.s[1];
      st.fold(op == MTL_EXPR_ADD ? x + y : (op == MTL_EXPR_SUB ? x - y : x * y));
    }
  }
  return st.s[0];
}

// The fact pass, two kernels:
//   k_fact_scan (one 4096-row tile per block, no barrier): load the
//     foreign keys and the first fact predicate's column together (dense
//     16-byte loads: rows are read once, in one round trip, instead of a
//     second dependent round for the rows that found a parent), probe the
//     parent bitmap (the parent key order of the facts keeps a wave's probes
//     within a few cache lines), evaluate the predicates, and write each
//     wave's hits (row, key slot) to that wave's own 1024-entry region of
//     the hit buffer with a wave-local prefix sum, plus the wave's count;
//   k_fact_acc (kAccTiles tiles per block): the block's hits are numbered
//     across its tiles' wave regions, one hit per thread: the slot rank
//     (prefix popcount of the parent bitmap), the expression columns and
//     the value are loaded and computed in parallel and added to the slot
//     with global atomics (hits are sparse and spread over consecutive
//     slots).  Each block writes the max |value| it added to tmax[block]
//     (no shared counter); the top-K merge reduces it for the overflow bound.
constexpr int kAccTiles = 16;
constexpr int kAccRegions = kAccTiles * FWAVES;
static_assert(kAccRegions <= WAVE, "region offsets are one wave's prefix sum");

// FAST (4-byte keys, host-checked min >= 0 and min + span <= 2^31): one
// 32-bit subtraction per key, since a negative key or one below min wraps to
// at least 2^31 - min >= span.
template <typename KT, bool FULL, bool FAST>
__device__ __forceinline__ void fact_keys(const FactArgs& a, uint64_t k0, uint32_t (&s)[FJ][4]) {
#pragma unroll
  for (int j = 0; j < FJ; ++j) {
    KT x[4];
    load4<KT, FULL>((const KT*)a.fk, k0 + (uint64_t)j * 256, a.n, a.fk_vec != 0, x);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      if constexpr (FAST) {
        const uint32_t o = (uint32_t)x[v] - (uint32_t)a.min;
        s[j][v] = o < (uint32_t)a.span ? o : 0xFFFFFFFFu;
      } else {
        uint64_t o = key_off((int64_t)x[v], a.min);
        // 0xFFFFFFFF marks "outside the parent key range" (span < 2^32)
        s[j][v] = o < a.span ? (uint32_t)o : 0xFFFFFFFFu;
      }
    }
  }
}

// PW: byte width of the first predicate's column (0 without predicates), so
// its values are held at their own width; bitmap words are read as 32-bit
// words.  Both keep the kernel under 64 VGPRs (8 waves per SIMD).
// FULL: the tile lies inside the table and its columns are aligned, so
// every column load is a straight-line vector load and all of them are in
// flight together.
// FAST: 32-bit key offsets and, for a 4-byte predicate, its int32 interval.
template <int PW, typename KT, bool FULL, bool FAST>
__device__ __forceinline__ void fact_scan_tile(const FactArgs& a) {
  const int tid = threadIdx.x, wave = tid / WAVE, lane = tid % WAVE;
  const uint64_t k0 = (uint64_t)blockIdx.x * FTILE + (uint64_t)wave * (FJ * 256) + (uint64_t)lane * 4;
  // first predicate's column, loaded with the keys
  using PT = typename std::conditional<
      PW == 8, uint64_t, typename std::conditional<PW == 1, uint8_t, uint32_t>::type>::type;
  PT px[FJ][4];
  if constexpr (PW != 0) {
#pragma unroll
    for (int j = 0; j < FJ; ++j)
      load4<PT, FULL>((const PT*)a.preds[0].col, k0 + (uint64_t)j * 256, a.n, a.pred0_vec != 0,
                      px[j]);
  }
  uint32_t s[FJ][4];
  fact_keys<KT, FULL, FAST>(a, k0, s);
  // rows in range and, when there is a predicate, passing it (its column
  // arrived with the keys): only these probe the parent bitmap
  uint32_t cand = 0;
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if ((FULL || k0 + (uint64_t)j * 256 + v < a.n) && s[j][v] != 0xFFFFFFFFu) {
        if constexpr (FAST && PW == 4) {
          if (!pred_pass_i32((uint32_t)px[j][v], a.preds[0])) continue;
        } else if constexpr (PW != 0) {
          if (!pred_pass(enc_raw((uint64_t)px[j][v], a.preds[0].kind), a.preds[0])) continue;
        }
        cand |= 1u << (4 * j + v);
      }
  uint32_t hits = 0;
  const uint32_t* bm32 = (const uint32_t*)a.bm;
  uint32_t far = cand;  // candidates still to probe one by one
  if (FULL) {
    // Clustered keys (a fact table ordered by its parent): a lane's four
    // consecutive rows of a step name one or two parents, so one two-word
    // window at the word of the step's first candidate covers them all: two
    // loads per step instead of four (the bitmap has a zero word past its
    // end).  Rows outside the window fall back to their own load below.
    uint32_t wb[FJ], lo[FJ], hi[FJ];
#pragma unroll
    for (int j = 0; j < FJ; ++j) {
      uint32_t s0 = 0;
#pragma unroll
      for (int v = 3; v >= 0; --v) s0 = ((cand >> (4 * j + v)) & 1u) ? s[j][v] : s0;
      wb[j] = s0 >> 5;
      lo[j] = ld_off32(bm32, wb[j] << 2);
      hi[j] = ld_off32(bm32, (wb[j] + 1) << 2);
    }
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const uint32_t b = 1u << (4 * j + v);
        const uint32_t d = s[j][v] - (wb[j] << 5);
        if ((cand & b) && d < 64) {
          far &= ~b;
          if (((d < 32 ? lo[j] : hi[j]) >> (d & 31)) & 1u) hits |= b;
        }
      }
  }
  if (__ballot(far != 0)) {  // wave-uniform
    uint32_t w[FJ][4];
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        w[j][v] = ((far >> (4 * j + v)) & 1u) ? ld_off32(bm32, (s[j][v] >> 5) << 2) : 0u;
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (((far >> (4 * j + v)) & 1u) && ((w[j][v] >> (s[j][v] & 31)) & 1u))
          hits |= 1u << (4 * j + v);
  }
  for (uint32_t p = 1; p < a.n_preds && hits; ++p) pred_rows(a.preds[p], k0, hits);
  // the wave's hits, lane by lane, to the wave's region.  Hits are sparse:
  // the loop runs once per hit of the lane (not 16 times with masked
  // stores), and a hit's slot is picked from s by a select chain (static
  // indices, so s stays in registers).
  const uint32_t c = (uint32_t)__popc(hits);
  const uint32_t incl = wave_inclusive_u32(c);
  const uint64_t region = (uint64_t)blockIdx.x * FTILE + (uint64_t)wave * (FJ * 256);
  uint32_t p = incl - c;
  uint32_t h = hits;
  while (h) {
    const int i = __ffs(h) - 1;
    h &= h - 1;
    uint32_t sv = s[0][0];
#pragma unroll
    for (int q = 1; q < FJ * 4; ++q) sv = i == q ? s[q >> 2][q & 3] : sv;
    a.hit_row[region + p] = (uint32_t)(k0 + (uint64_t)(i >> 2) * 256 + (i & 3));
    a.hit_slot[region + p] = sv;
    ++p;
  }
  if (lane == WAVE - 1) a.hit_count[(uint64_t)blockIdx.x * FWAVES + wave] = incl;
}

template <int PW, typename KT, bool FAST>
__global__ __launch_bounds__(FBLOCK, 8) void k_fact_scan(FactArgs a) {
  const bool full = a.fk_vec && (PW == 0 || a.pred0_vec) &&
                    (uint64_t)(blockIdx.x + 1) * FTILE <= a.n;  // block-uniform
  if (full) fact_scan_tile<PW, KT, true, FAST>(a);
  else fact_scan_tile<PW, KT, false, FAST>(a);
}

template <bool DBL>
__global__ __launch_bounds__(FBLOCK) void k_fact_acc(FactArgs a) {
  __shared__ uint32_t s_off[kAccRegions + 1];
  const int tid = threadIdx.x;
  const uint64_t r0 = (uint64_t)blockIdx.x * kAccRegions;  // first wave region
  if (tid < WAVE) {
    const uint64_t r = r0 + tid;
    const uint32_t cnt = (tid < kAccRegions && r < a.n_regions) ? a.hit_count[r] : 0u;
    const uint32_t incl = wave_inclusive_u32(cnt);
    if (tid < kAccRegions) s_off[tid] = incl - cnt;
    if (tid == kAccRegions - 1) s_off[kAccRegions] = incl;
  }
  __syncthreads();
  const uint32_t total = s_off[kAccRegions];
  bool ovf = false;
  uint64_t mx = 0;
  using V = typename std::conditional<DBL, double, int64_t>::type;
  // block-uniform trip count: every lane takes part in the wave reduction
  for (uint32_t e0 = 0; e0 < total; e0 += FBLOCK) {
    const uint32_t e = e0 + tid;
    const bool live = e < total;
    uint32_t r = 0xFFFFFFFFu;
    V val = 0;
    if (live) {
      // last region whose start is <= e (empty regions share a start)
      uint32_t q = 0, hi = kAccRegions - 1;
      while (q < hi) {
        const uint32_t mid = (q + hi + 1) >> 1;
        if (s_off[mid] <= e) q = mid;
        else hi = mid - 1;
      }
      const uint64_t at = (r0 + q) * (FJ * 256) + (e - s_off[q]);
      const uint32_t row = a.hit_row[at], sv = a.hit_slot[at];
      V cv[MTL_EXPR_MAX_COLS];
      uint64_t w;
      uint32_t rk;
      if (a.expr_w8) {
        // every column 8 bytes wide: the column loads and the slot lookup go
        // out together, one round trip, converted afterwards
        uint64_t raw[MTL_EXPR_MAX_COLS];
#pragma unroll
        for (int c = 0; c < MTL_EXPR_MAX_COLS; ++c) raw[c] = ((const uint64_t*)a.expr.col[c])[row];
        w = a.bm[sv >> 6];
        rk = a.rank[sv >> 6];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int c = 0; c < MTL_EXPR_MAX_COLS; ++c) {
          if constexpr (DBL)
            cv[c] = a.expr.col_type[c] == MTL_DOUBLE ? __longlong_as_double((long long)raw[c])
                                                     : (double)(int64_t)raw[c];
          else cv[c] = (int64_t)raw[c];
          if ((uint32_t)c >= a.expr.n_cols) cv[c] = 0;
        }
      } else {
#pragma unroll
        for (int c = 0; c < MTL_EXPR_MAX_COLS; ++c)
          if ((uint32_t)c < a.expr.n_cols) {
            if constexpr (DBL) cv[c] = load_f64(a.expr.col[c], a.expr.col_type[c], row);
            else cv[c] = load_key(a.expr.col[c], a.expr.col_type[c], row);
          } else {
            cv[c] = 0;
          }
        w = a.bm[sv >> 6];
        rk = a.rank[sv >> 6];
      }
      const uint64_t below = (sv & 63) ? (w & ((1ull << (sv & 63)) - 1)) : 0ull;
      r = rk + (uint32_t)__popcll(below);
      if constexpr (DBL) {
        val = expr_f64_vals(a.expr, cv);
      } else {
        val = expr_i64_vals(a.expr, cv, ovf);
        const uint64_t av = val < 0 ? (uint64_t)0 - (uint64_t)val : (uint64_t)val;
        mx = av > mx ? av : mx;
      }
    }
    // Hits of one parent row sit in consecutive lanes (the hit lists keep
    // row order): sum each run of equal slots across the wave and let only
    // the run's first lane update the slot, so a slot gets one atomic per
    // run instead of one per hit (same-address atomics serialise).
    const int lane = tid % WAVE;
    const uint32_t prev = __shfl_up(r, 1, WAVE);
    const uint32_t head = (lane == 0 || prev != r) ? 1u : 0u;
    const uint32_t seg = wave_inclusive_u32(head);
    uint32_t n = live ? 1u : 0u;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
      const uint32_t os = __shfl_down(seg, off, WAVE);
      const uint32_t on = __shfl_down(n, off, WAVE);
      V ov;
      if constexpr (DBL) {
        ov = __shfl_down(val, off, WAVE);
      } else {
        ov = (int64_t)__shfl_down((unsigned long long)val, off, WAVE);
      }
      if (lane + off < WAVE && os == seg) {
        if constexpr (DBL) val += ov;
        else val = (int64_t)((uint64_t)val + (uint64_t)ov);  // wrapping, as the atomics
        n += on;
      }
    }
    if (head && r != 0xFFFFFFFFu) {
      if constexpr (DBL) unsafeAtomicAdd((double*)&a.acc[r], val);
      else atomicAdd((unsigned long long*)&a.acc[r], (unsigned long long)val);
      atomicAdd(&a.cnt[r], n);
    }
  }
  if (ovf) atomicOr(a.err, 16u);
  if (!DBL) {
    mx = block_max_u64(mx);
    if (tid == 0) a.tmax[blockIdx.x] = mx;
  }
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
  TopkStage stage;    // full keys of every wave-list entry
  TopkLists lists0;   // per-block lists written by phase 1
  mtl_star_row* out;
  uint32_t* out_n;  // [0] = rows, [1] = error bits copied from err
  uint32_t* err;
  const unsigned long long* tmax;  // per fact tile: max |value| added
  uint32_t n_tmax;
  uint32_t* bcmax;                 // per top-K block: max count of its slots
  uint32_t n_bcmax;
  uint64_t p_rows;  // row_of entries are checked against it before use
  uint32_t grid;    // phase-1 blocks (shared multi-query launches)
};

// Tie-break words of slot r whose parent row is `row` (read with the slot).
__device__ __forceinline__ void star_keys(const TopkArgs& a, uint32_t row, uint64_t k[4]) {
  if (row >= a.p_rows) {  // cannot happen for consistent inputs; never read past P
    atomicOr(a.err, 64u);
    row = 0;
  }
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

template <int R>
__device__ __forceinline__ void star_topk_body(const TopkArgs& a, uint32_t bid, uint32_t nb) {
  using S = TopkShape<R>;
  __shared__ TopkBlockLds<S::W, S::CAP> sl;
  topk_block_init(sl);
  WaveTopk<R> wl;
  wl.init(a.k);
  const uint32_t m = *a.m_count;
  const uint32_t chunk = (m + nb - 1) / nb;
  const uint32_t start = bid * chunk;
  const uint32_t end = min(m, start + chunk);
  constexpr int E = 8;  // entries per thread per pass, loaded together
  uint32_t cmax = 0;
  for (uint32_t base = start; base < end; base += S::NT * E) {
    uint32_t cv[E], rv[E];
    uint64_t av[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      uint32_t r = base + e * S::NT + threadIdx.x;
      cv[e] = r < end ? a.cnt[r] : 0u;
      av[e] = r < end ? a.acc[r] : 0ull;
      rv[e] = r < end ? a.row_of[r] : 0u;
    }
#pragma unroll
    for (int e = 0; e < E; ++e) cmax = cv[e] > cmax ? cv[e] : cmax;
    if (R <= 2 && base == start) {  // seed the block bound from the first round
      uint64_t m1 = ~0ull, m2 = ~0ull;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const uint64_t k0 = a.is_double ? ~enc_f64(__longlong_as_double((long long)av[e]))
                                        : ~enc_i64((int64_t)av[e]);
        if (cv[e] == 0) continue;
        if (R == 1) m1 = k0 < m1 ? k0 : m1;  // K <= 64 needs the minimum only
        else topk_min2(k0, m1, m2);
      }
      topk_seed_thr(m1, m2, a.k, sl);
    }
    const uint64_t bthr = __hip_atomic_load(&sl.thr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const uint32_t r = base + e * S::NT + threadIdx.x;
      uint64_t k[4] = {0, 0, 0, 0};
      k[0] = a.is_double ? ~enc_f64(__longlong_as_double((long long)av[e]))
                         : ~enc_i64((int64_t)av[e]);
      const bool pre = cv[e] != 0 && k[0] <= wl.thr0 && k[0] <= bthr;
      if (pre) star_keys(a, rv[e], k);
      wl.offer(pre, k, r);
    }
    topk_share_thr(wl, sl);
  }
  topk_block_finish<S::NT>(wl, sl, a.stage, a.lists0, bid);
  cmax = block_max_u32(cmax);
  if (threadIdx.x == 0) a.bcmax[bid] = cmax;
}

template <int R>
__global__ __launch_bounds__(TopkShape<R>::NT) void k_star_topk(TopkArgs a) {
  star_topk_body<R>(a, blockIdx.x, gridDim.x);
}

// Phase 1 of a batch of queries' top-K: blockIdx.y is the query.
template <int R>
__global__ __launch_bounds__(TopkShape<R>::NT) void k_mq_star_topk(const TopkArgs* qs) {
  const TopkArgs& a = qs[blockIdx.y];
  if (blockIdx.x >= a.grid) return;
  star_topk_body<R>(a, blockIdx.x, a.grid);
}

// Merge level: block b merges lists [b*F, b*F+F) of `in`; the single block
// of the final level also checks the overflow bound and writes the rows.
__device__ __forceinline__ void star_merge_body(const TopkArgs& a, const TopkLists& in,
                                                const TopkLists& out, uint32_t L,
                                                uint32_t final_level, uint32_t bid) {
  __shared__ TopkMergeLds s;
  // the overflow bound's inputs are independent of the merge: on the final
  // level their loads go out first and arrive while the lists merge
  uint64_t mx = 0;
  uint32_t cm = 0;
  if (final_level && !a.is_double) {
    // eight independent loads in flight per thread per round
    for (uint32_t i0 = 0; i0 < a.n_tmax; i0 += TOPK_MERGE_THREADS * 8) {
      uint64_t t[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const uint32_t i = i0 + u * TOPK_MERGE_THREADS + threadIdx.x;
        t[u] = i < a.n_tmax ? a.tmax[i] : 0ull;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) mx = t[u] > mx ? t[u] : mx;
    }
    for (uint32_t i = threadIdx.x; i < a.n_bcmax; i += TOPK_MERGE_THREADS)
      cm = a.bcmax[i] > cm ? a.bcmax[i] : cm;
  }
  const int cur = topk_merge_lists(s, in, L, a.k, a.stage, bid);
  if (!final_level) {
    topk_write_list(s, cur, out, a.k, bid);
    return;
  }
  // |SUM| <= count x max|value| for every group: flag the run when that
  // bound exceeds int64 (an int64 accumulator could have wrapped)
  if (!a.is_double) {
    mx = block_max_u64(mx);
    cm = block_max_u32(cm);
    unsigned long long prod;
    if (threadIdx.x == 0 &&
        (__builtin_mul_overflow((unsigned long long)cm, (unsigned long long)mx, &prod) ||
         prod > 0x7FFFFFFFFFFFFFFFull))
      atomicOr(a.err, 32u);
  }
  const uint32_t nn = s.n[cur][0];
  for (uint32_t i = threadIdx.x; i < nn; i += TOPK_MERGE_THREADS) {
    const uint32_t r = a.stage.cid[s.g[cur][i]];
    mtl_star_row row;
    row.p_row = a.row_of[r];
    row.count = a.cnt[r];
    uint64_t v = a.acc[r];
    row.sum_i64 = a.is_double ? 0 : (int64_t)v;
    row.sum_f64 = a.is_double ? __longlong_as_double((long long)v) : (double)(int64_t)v;
    a.out[i] = row;
  }
  __syncthreads();  // every thread's atomicOr above precedes the error read
  if (threadIdx.x == 0) {
    a.out_n[0] = nn;
    a.out_n[1] = __hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__global__ __launch_bounds__(TOPK_MERGE_THREADS) void k_star_topk_merge(TopkArgs a, TopkLists in,
                                                                        TopkLists out, uint32_t L,
                                                                        uint32_t final_level) {
  star_merge_body(a, in, out, L, final_level, blockIdx.x);
}

// One merge level of a batch of queries: blockIdx.y selects the query's
// level (its top-K arguments, input and output lists, list count).
struct MQMergeQ {
  const TopkArgs* ta;
  TopkLists in, out;
  uint32_t L, final_level, blocks;
};

__global__ __launch_bounds__(TOPK_MERGE_THREADS) void k_mq_star_merge(const MQMergeQ* qs) {
  const MQMergeQ& m = qs[blockIdx.y];
  if (blockIdx.x >= m.blocks) return;
  star_merge_body(*m.ta, m.in, m.out, m.L, m.final_level, blockIdx.x);
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
  if (sp->pspan == 0 || sp->pspan >= (1ull << 32))
    return set_error(MTL_UNSUPPORTED, "parent key span %llu too wide for a bitmap",
                     (unsigned long long)sp->pspan);
  sp->pwords = ceil_div(sp->pspan, 64);
  if (d->d_rows) {
    if (d->d_key_max < d->d_key_min) return set_error(MTL_ERR_INVALID_ARGUMENT, "dim key range");
    sp->dspan = key_off(d->d_key_max, d->d_key_min) + 1;
    if (sp->dspan == 0 || sp->dspan >= (1ull << 32))
      return set_error(MTL_UNSUPPORTED, "dimension key span too wide for a bitmap");
    sp->dwords = ceil_div(sp->dspan, 64);
  }
  // about one round of 8 slots per thread per block (the slot count is
  // at most P): 1024-thread blocks, one per CU, for K <= 128; 256-thread
  // blocks, up to three per CU, above
  const uint64_t nt = d->k <= 128 ? 1024 : 256;
  uint64_t g = ceil_div(d->p_rows ? d->p_rows : 1, nt * 8);
  sp->topk_grid = (uint32_t)std::max<uint64_t>(
      1, std::min<uint64_t>(g, (d->k <= 128 ? 1 : 3) * (uint64_t)cu_count()));
  // when one merge block can take every phase-1 list while keeping at least
  // half the CUs busy in phase 1, a single merge launch suffices
  const uint32_t fan = topk_fan(d->k);
  if (sp->topk_grid > fan && 2 * fan >= (uint32_t)cu_count()) sp->topk_grid = fan;
  return MTL_OK;
}

// One star query's device scratch (star_enqueue and the shared batch).
struct StarBufs {
  void *dim_bm = nullptr, *dbits = nullptr, *par_bm, *par_rank, *pbits, *row_of, *acc, *cnt,
       *misc;
  void *st_r, *ck[4], *cid, *lk0[2], *lg[2], *ln[2], *bcmax, *tmax;
  uint32_t rtiles = 0;
  uint32_t n_tmax = 0;  // entries of tmax (the overflow bound's per-block maxima)
};

// Arguments of one query's rank kernel (k_rank's parameters).
struct RankLaunch {
  const uint64_t* bm;
  uint64_t words;
  uint32_t rtiles;
  uint32_t* rank;
  uint64_t* status;
  uint32_t* ctr;
  uint32_t* err;
  uint32_t* par_total;
  PopArgs pa;
  uint32_t grid;
};

__global__ __launch_bounds__(RBLOCK) void k_mq_rank(const RankLaunch* qs) {
  const RankLaunch& a = qs[blockIdx.y];
  if (blockIdx.x >= a.grid) return;
  rank_body(a.bm, a.words, a.rtiles, a.rank, a.status, a.ctr, a.err, a.par_total, a.pa,
            blockIdx.x, a.grid);
}

static RankLaunch make_rank(const mtl_star_desc* d, const StarPlan& sp, const StarBufs& b) {
  const uint32_t ptiles = filter_tiles(d->p_rows);
  uint32_t* mu = (uint32_t*)b.misc;
  RankLaunch r{};
  r.pa.pbits = (const uint64_t*)b.pbits;
  r.pa.pbits_words = (uint64_t)ptiles * kTileRows / 64;
  if (d->d_rows) {
    r.pa.dim_bm = (const uint64_t*)b.dim_bm;
    r.pa.dim_words = sp.dwords;
    r.pa.dbits = (const uint64_t*)b.dbits;
    r.pa.dbits_words = (uint64_t)filter_tiles(d->d_rows) * kTileRows / 64;
  }
  r.pa.pops = mu + 7;  // [7] P pass bits, [8] D bitmap, [9] D pass bits
  const uint32_t rt = std::max<uint32_t>(b.rtiles, 1);
  const uint64_t pop_words =
      std::max(r.pa.pbits_words, std::max(r.pa.dim_words, r.pa.dbits_words));
  const uint32_t pop_blocks =
      (uint32_t)std::min<uint64_t>(ceil_div(pop_words, (uint64_t)RBLOCK * 8), 2048);
  r.bm = (const uint64_t*)b.par_bm;
  r.words = sp.pwords;
  r.rtiles = rt;
  r.rank = (uint32_t*)b.par_rank;
  r.status = (uint64_t*)b.st_r;
  r.ctr = mu + 2;
  r.err = mu + 3;
  r.par_total = mu + 6;
  r.grid = std::max(rt, pop_blocks);
  return r;
}

// k_map4: 4-byte parent keys aligned for vector loads.
static bool map_k4(const mtl_star_desc* d) {
  return type_width(d->p_key_type) == 4 && d->p_key_type != MTL_FLOAT &&
         ((uintptr_t)d->p_key % 16) == 0;
}

static MapArgs make_map(const mtl_star_desc* d, const StarPlan& sp, const StarBufs& b) {
  uint32_t* mu = (uint32_t*)b.misc;
  MapArgs ma{};
  ma.sel_bm = (const uint64_t*)b.pbits;
  ma.n_rows = d->p_rows;
  ma.key = d->p_key;
  ma.key_type = d->p_key_type;
  ma.min = d->p_key_min;
  ma.span = sp.pspan;
  ma.bm = (const uint64_t*)b.par_bm;
  ma.rank = (const uint32_t*)b.par_rank;
  ma.row_of = (uint32_t*)b.row_of;
  ma.acc = (uint64_t*)b.acc;
  ma.cnt = (uint32_t*)b.cnt;
  ma.par_total = mu + 6;
  ma.pops = mu + 7;
  ma.has_dim = d->d_rows ? 1u : 0u;
  ma.err = mu + 3;
  ma.grid = (uint32_t)std::max<uint64_t>(
      1, std::min<uint64_t>(ceil_div(d->p_rows ? d->p_rows : 1, 256 * 16), 4096));
  return ma;
}

// Steps 3 and 4: prefix popcounts of the parent bitmap, then the slot map
// (row_of) and the cleared accumulators.
static int32_t launch_rank_map(const mtl_star_desc* d, const StarPlan& sp, const StarBufs& b,
                               mtl_stream* s) {
  hipStream_t st = s->stream;
  {
    ProfScope ps(s, "k_rank");
    const RankLaunch r = make_rank(d, sp, b);
    hipLaunchKernelGGL(k_rank, dim3(r.grid), dim3(RBLOCK), 0, st, r.bm, r.words, r.rtiles,
                       r.rank, r.status, r.ctr, r.err, r.par_total, r.pa);
  }
  MTL_HIP(hipGetLastError());
  const MapArgs ma = make_map(d, sp, b);
  ProfScope ps(s, "k_map");
  if (map_k4(d)) hipLaunchKernelGGL(k_map4, dim3(ma.grid), dim3(256), 0, st, ma);
  else hipLaunchKernelGGL(k_map, dim3(ma.grid), dim3(256), 0, st, ma);
  MTL_HIP(hipGetLastError());
  return MTL_OK;
}

static TopkArgs make_topk(const mtl_star_desc* d, const StarPlan& sp, const StarBufs& b,
                          mtl_star_row* out, uint32_t* out_n) {
  uint32_t* mu = (uint32_t*)b.misc;
  TopkArgs ta{};
  ta.m_count = mu + 6;
  ta.acc = (const uint64_t*)b.acc;
  ta.cnt = (const uint32_t*)b.cnt;
  ta.row_of = (const uint32_t*)b.row_of;
  ta.is_double = d->agg.is_double;
  ta.k = d->k;
  ta.n_ties = d->n_ties;
  for (uint32_t t = 0; t < d->n_ties; ++t) {
    ta.tie_col[t] = d->tie_col[t];
    ta.tie_type[t] = d->tie_type[t];
    ta.tie_desc[t] = d->tie_desc[t];
  }
  for (int f = 0; f < 4; ++f) ta.stage.ck[f] = (uint64_t*)b.ck[f];
  ta.stage.cid = (uint32_t*)b.cid;
  ta.lists0 = TopkLists{(uint64_t*)b.lk0[0], (uint32_t*)b.lg[0], (uint32_t*)b.ln[0]};
  ta.out = out;
  ta.out_n = out_n;
  ta.err = mu + 3;
  ta.tmax = (const unsigned long long*)b.tmax;
  ta.n_tmax = d->agg.is_double ? 0u : b.n_tmax;
  ta.bcmax = (uint32_t*)b.bcmax;
  ta.n_bcmax = sp.topk_grid;
  ta.p_rows = d->p_rows;
  ta.grid = sp.topk_grid;
  return ta;
}

// Step 6: top-K over the accumulators and the output rows.
static int32_t launch_star_topk(const mtl_star_desc* d, const StarPlan& sp, const StarBufs& b,
                                mtl_star_row* out, uint32_t* out_n, mtl_stream* s) {
  hipStream_t st = s->stream;
  const TopkArgs ta = make_topk(d, sp, b, out, out_n);
  TopkLists lists[2];
  for (int i = 0; i < 2; ++i)
    lists[i] = TopkLists{(uint64_t*)b.lk0[i], (uint32_t*)b.lg[i], (uint32_t*)b.ln[i]};
  {
    ProfScope ps(s, "k_star_topk");
    if (d->k <= 64)
      hipLaunchKernelGGL(k_star_topk<1>, dim3(sp.topk_grid), dim3(TopkShape<1>::NT), 0, st, ta);
    else if (d->k <= 128)
      hipLaunchKernelGGL(k_star_topk<2>, dim3(sp.topk_grid), dim3(TopkShape<2>::NT), 0, st, ta);
    else
      hipLaunchKernelGGL(k_star_topk<8>, dim3(sp.topk_grid), dim3(TopkShape<8>::NT), 0, st, ta);
  }
  MTL_HIP(hipGetLastError());
  // merge levels: F lists per block until one list remains; the last
  // level (one block) writes the rows
  ProfScope ps2(s, "k_star_topk_merge");
  const uint32_t fan = topk_fan(d->k);
  uint32_t L = sp.topk_grid;
  int in = 0;
  while (true) {
    const uint32_t blocks = (uint32_t)ceil_div(L, fan);
    const uint32_t fin = blocks == 1 ? 1u : 0u;
    hipLaunchKernelGGL(k_star_topk_merge, dim3(blocks), dim3(TOPK_MERGE_THREADS), 0, st, ta,
                       lists[in], lists[in ^ 1], L, fin);
    MTL_HIP(hipGetLastError());
    if (fin) break;
    L = blocks;
    in ^= 1;
  }
  return MTL_OK;
}

// Top-K scratch of one query (staged wave-list entries, two ping-pong sets
// of lists, per-block count maxima) from the arena.
static int32_t alloc_topk_bufs(const mtl_star_desc* d, const StarPlan& sp, Arena& ar,
                               StarBufs& b) {
  const uint64_t tk_waves = d->k <= 128 ? 16 : 4;
  const uint64_t stage_n = (uint64_t)sp.topk_grid * tk_waves * d->k;
  for (int f = 0; f < 4; ++f) MTL_TRY(ar.alloc(stage_n * 8, &b.ck[f]));
  MTL_TRY(ar.alloc(stage_n * 4, &b.cid));
  const uint64_t la = sp.topk_grid, lb = ceil_div(sp.topk_grid, topk_fan(d->k));
  MTL_TRY(ar.alloc(la * d->k * 8, &b.lk0[0]));
  MTL_TRY(ar.alloc(la * d->k * 4, &b.lg[0]));
  MTL_TRY(ar.alloc(la * 4, &b.ln[0]));
  MTL_TRY(ar.alloc(lb * d->k * 8, &b.lk0[1]));
  MTL_TRY(ar.alloc(lb * d->k * 4, &b.lg[1]));
  MTL_TRY(ar.alloc(lb * 4, &b.ln[1]));
  MTL_TRY(ar.alloc((uint64_t)sp.topk_grid * 4, &b.bcmax));
  return MTL_OK;
}

// Enqueue the whole pipeline on s.  out (device, k rows) and out_n (device,
// 2 x uint32: rows, error bits) receive the result.
static int32_t star_enqueue(const mtl_star_desc* d, const StarPlan& sp, mtl_stream* s,
                            mtl_star_row* out, uint32_t* out_n) {
  hipStream_t st = s->stream;
  Arena& ar = s->arena;
  const uint64_t prows = d->p_rows;
  StarBufs b;
  if (d->d_rows) {
    MTL_TRY(ar.alloc(sp.dwords * 8, &b.dim_bm));
    MTL_TRY(ar.alloc((uint64_t)filter_tiles(d->d_rows) * kTileRows / 8, &b.dbits));
  }
  // one zero word past the end: the fact scan reads two-word windows
  MTL_TRY(ar.alloc((sp.pwords + 1) * 8, &b.par_bm));
  MTL_TRY(ar.alloc(sp.pwords * 4, &b.par_rank));
  const uint32_t ptiles = filter_tiles(prows);
  MTL_TRY(ar.alloc((uint64_t)ptiles * kTileRows / 8, &b.pbits));
  MTL_TRY(ar.alloc((prows ? prows : 1) * 4, &b.row_of));
  MTL_TRY(ar.alloc((prows ? prows : 1) * 8, &b.acc));
  MTL_TRY(ar.alloc((prows ? prows : 1) * 4, &b.cnt));
  // misc words: [2] rank ticket, [3] error bits,
  // [6] parent bitmap popcount, [7] P pass-bit popcount, [8] dimension
  // bitmap popcount, [9] D pass-bit popcount
  MTL_TRY(ar.alloc(256, &b.misc));
  b.rtiles = (uint32_t)ceil_div(sp.pwords, RTILE);
  MTL_TRY(ar.alloc((uint64_t)std::max<uint32_t>(b.rtiles, 1) * 8, &b.st_r));
  MTL_TRY(alloc_topk_bufs(d, sp, ar, b));
  const uint32_t ftiles = d->f_rows ? filter_tiles(d->f_rows) : 0u;
  const uint32_t acc_blocks = (uint32_t)ceil_div(ftiles, kAccTiles);
  MTL_TRY(ar.alloc((uint64_t)std::max<uint32_t>(acc_blocks, 1) * 8, &b.tmax));
  b.n_tmax = acc_blocks;
  uint32_t* mu = (uint32_t*)b.misc;

  ZeroList zl;
  if (b.dim_bm) zl.add((uint32_t*)b.dim_bm, sp.dwords * 2);
  zl.add((uint32_t*)b.par_bm, (sp.pwords + 1) * 2);
  zl.add((uint32_t*)b.misc, 64);
  zl.add((uint32_t*)b.st_r, (uint64_t)std::max<uint32_t>(b.rtiles, 1) * 2);
  MTL_TRY(launch_zero(zl, st));

  // 1. dimension bitmap
  if (d->d_rows) {
    // mark-only filter pass: vectorised predicates, combined bitmap atomics
    FilterArgs fa{};
    for (uint32_t p = 0; p < d->d_n_preds; ++p) fa.preds[p] = compile_pred(d->d_preds[p]);
    fa.n_preds = d->d_n_preds;
    fa.n_rows = d->d_rows;
    fa.err = mu + 3;
    fa.mark_bm = (uint64_t*)b.dim_bm;
    fa.mark_key = d->d_key;
    fa.mark_key_type = d->d_key_type;
    fa.mark_min = d->d_key_min;
    fa.mark_span = sp.dspan;
    fa.sel_bm = (uint64_t*)b.dbits;
    ProfScope ps(s, "k_mark_dim");
    MTL_TRY(launch_filter(fa, d->d_rows, st));
  }
  // 2. parent filter + semi-join probe + compaction + parent bitmap
  {
    FilterArgs fa{};
    for (uint32_t p = 0; p < d->p_n_preds; ++p) fa.preds[p] = compile_pred(d->p_preds[p]);
    fa.n_preds = d->p_n_preds;
    fa.n_rows = prows;
    fa.sel_bm = (uint64_t*)b.pbits;
    fa.err = mu + 3;
    if (d->d_rows) {
      fa.probe_bm = (const uint64_t*)b.dim_bm;
      fa.probe_key = d->p_fk;
      fa.probe_key_type = d->p_fk_type;
      fa.probe_min = d->d_key_min;
      fa.probe_span = sp.dspan;
    }
    fa.mark_bm = (uint64_t*)b.par_bm;
    fa.mark_key = d->p_key;
    fa.mark_key_type = d->p_key_type;
    fa.mark_min = d->p_key_min;
    fa.mark_span = sp.pspan;
    ProfScope ps(s, "k_filter_parent");
    MTL_TRY(launch_filter(fa, prows, st));
  }
  // 3-4. rank, slot map
  MTL_TRY(launch_rank_map(d, sp, b, s));
  // 5. fact pass
  if (d->f_rows) {
    FactArgs fa{};
    fa.n = d->f_rows;
    fa.fk = d->f_fk;
    fa.fk_type = d->f_fk_type;
    fa.min = d->p_key_min;
    fa.span = sp.pspan;
    fa.bm = (const uint64_t*)b.par_bm;
    fa.rank = (const uint32_t*)b.par_rank;
    for (uint32_t p = 0; p < d->f_n_preds; ++p) fa.preds[p] = compile_pred(d->f_preds[p]);
    fa.n_preds = d->f_n_preds;
    fa.expr = d->agg;
    fa.expr_w8 = d->agg.n_cols > 0;
    for (uint32_t c = 0; c < d->agg.n_cols && c < MTL_EXPR_MAX_COLS; ++c)
      fa.expr_w8 &= type_width(d->agg.col_type[c]) == 8 ? 1u : 0u;
    if (fa.expr_w8)
      for (uint32_t c = d->agg.n_cols; c < MTL_EXPR_MAX_COLS; ++c) {
        fa.expr.col[c] = d->agg.col[0];
        fa.expr.col_type[c] = d->agg.col_type[0];
      }
    fa.acc = (uint64_t*)b.acc;
    fa.cnt = (uint32_t*)b.cnt;
    fa.err = mu + 3;
    fa.fk_vec = ((uintptr_t)d->f_fk % 16) == 0;
    fa.tmax = (unsigned long long*)b.tmax;
    fa.pred0_vec = d->f_n_preds && ((uintptr_t)d->f_preds[0].column % 16) == 0;
    void *hrow, *hslot, *hcnt;
    MTL_TRY(ar.alloc((uint64_t)ftiles * FTILE * 4, &hrow));
    MTL_TRY(ar.alloc((uint64_t)ftiles * FTILE * 4, &hslot));
    MTL_TRY(ar.alloc((uint64_t)ftiles * FWAVES * 4, &hcnt));
    fa.hit_row = (uint32_t*)hrow;
    fa.hit_slot = (uint32_t*)hslot;
    fa.hit_count = (uint32_t*)hcnt;
    fa.n_regions = (uint64_t)ftiles * FWAVES;
    {
      ProfScope ps(s, "k_fact_scan");
      const uint32_t pw = d->f_n_preds ? pred_width(fa.preds[0].kind) : 0u;
      const int kw = d->f_fk_type == MTL_DICT8 ? 1 : type_width(d->f_fk_type) == 8 ? 8 : 4;
      // 4-byte keys in [0, 2^31) and an integer 4-byte predicate (or none):
      // the 32-bit instantiation (32-bit key offsets and predicate compares;
      // 58 VGPRs without spills against 64 with)
      const bool fast = kw == 4 && d->p_key_min >= 0 &&
                        (uint64_t)d->p_key_min + sp.pspan <= (1ull << 31) &&
                        (pw == 0 || (pw == 4 && fa.preds[0].kind == PK_I32));
#define MTL_FACT_SCAN(P, K) \
  hipLaunchKernelGGL((k_fact_scan<P, K, false>), dim3(ftiles), dim3(FBLOCK), 0, st, fa)
#define MTL_FACT_SCAN_K(P)                  \
  if (kw == 4) MTL_FACT_SCAN(P, int32_t);   \
  else if (kw == 8) MTL_FACT_SCAN(P, int64_t); \
  else MTL_FACT_SCAN(P, uint8_t)
      if (fast && pw == 4) {
        hipLaunchKernelGGL((k_fact_scan<4, int32_t, true>), dim3(ftiles), dim3(FBLOCK), 0, st, fa);
      } else if (fast) {
        hipLaunchKernelGGL((k_fact_scan<0, int32_t, true>), dim3(ftiles), dim3(FBLOCK), 0, st, fa);
      } else if (pw == 0) { MTL_FACT_SCAN_K(0); }
      else if (pw == 1) { MTL_FACT_SCAN_K(1); }
      else if (pw == 4) { MTL_FACT_SCAN_K(4); }
      else { MTL_FACT_SCAN_K(8); }
#undef MTL_FACT_SCAN_K
#undef MTL_FACT_SCAN
    }
    MTL_HIP(hipGetLastError());
    ProfScope ps(s, "k_fact_acc");
    if (d->agg.is_double)
      hipLaunchKernelGGL(k_fact_acc<true>, dim3(acc_blocks), dim3(FBLOCK), 0, st, fa);
    else
      hipLaunchKernelGGL(k_fact_acc<false>, dim3(acc_blocks), dim3(FBLOCK), 0, st, fa);
    MTL_HIP(hipGetLastError());
  }
  // 6. top-K + output rows
  return launch_star_topk(d, sp, b, out, out_n, s);
}

static int32_t star_error(uint32_t err) {
  if (err & 1u) return set_error(MTL_ERR_PROCESSING, "processing error: look-back wait exceeded");
  if (err & 64u) return set_error(MTL_ERR_PROCESSING, "processing error: inconsistent group map");
  if (err & 8u)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "a key lies outside its stated [min, max]");
  if (err & (2u | 4u))
    return set_error(MTL_UNSUPPORTED, "join build key is not unique (fallback)");
  if (err & 16u) return set_error(MTL_ERR_OVERFLOW, "Overflow in aggregate expression");
  if (err & 32u) return set_error(MTL_UNSUPPORTED, "int64 SUM overflow (needs HUGEINT)");
  return MTL_OK;
}

int32_t star_check_error(uint32_t err) { return star_error(err); }

// One integer expression value per row (an aggregate over an expression of
// joined columns, gathered into row order first).
__global__ __launch_bounds__(256) void k_expr_i64(mtl_expr e, uint64_t n, int64_t* out,
                                                  uint32_t* err) {
  bool ovf = false;
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    int64_t cv[MTL_EXPR_MAX_COLS];
#pragma unroll
    for (int c = 0; c < MTL_EXPR_MAX_COLS; ++c)
      cv[c] = (uint32_t)c < e.n_cols ? load_key(e.col[c], e.col_type[c], i) : 0;
    out[i] = expr_i64_vals(e, cv, ovf);
  }
  if (__any(ovf) && (threadIdx.x & 63) == 0) atomicOr(err, 1u);
}

int32_t expr_eval_i64(const mtl_expr& e, uint64_t n, int64_t* out, uint32_t* err,
                      hipStream_t st) {
  if (e.is_double) return set_error(MTL_UNSUPPORTED, "floating-point aggregate expression");
  MTL_TRY(check_expr(e));
  if (!n) return MTL_OK;
  const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_expr_i64, dim3((uint32_t)blocks), dim3(256), 0, st, e, n, out, err);
  MTL_HIP(hipGetLastError());
  return MTL_OK;
}

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
  MTL_HIP(stream_wait(s->stream));
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

// ---------------------------------------------------------------------------
// Shared passes of a batch of star queries (star.hpp).  Each of the three
// table passes is one launch for the whole batch with one block per tile of
// the table's largest row count in the batch: the block loads the tile's
// key and predicate columns once and then applies, query by query, each
// covering query's predicate and row count, writing that query's pass bits
// and bitmap marks.  Queries with the same dimension predicate share one
// dimension slot (bitmap and pass bits), built once and probed once per
// parent tile.  The fact pass adds each hit straight into its query's
// accumulator (rank, expression columns, value): there are no hit lists, so
// a batch needs no scratch proportional to the fact table.  The per-query
// steps between the passes (ranks, slot map, top-K and its merge levels)
// are one launch each for the batch, blockIdx.y selecting the query.
// ---------------------------------------------------------------------------
namespace mtl {

struct MQStarQ {
  uint64_t d_rows, p_rows, f_rows;
  PredC dpred, ppred, fpred;
  uint32_t dslot;  // dimension slot
  uint64_t pspan;  // parent key span of this query
  uint64_t* dim_bm;
  uint64_t* dbits;
  uint64_t* pbits;
  uint64_t* par_bm;
  const uint32_t* rank;
  uint64_t* acc;
  uint32_t* cnt;
  uint32_t* err;
  unsigned long long* tmax;  // max |value| added (overflow bound)
};

// A dimension predicate shared by the queries that use it: its rows, key
// bitmap, pass bits and error word.
struct MQDimSlot {
  PredC pred;
  uint64_t rows;
  uint64_t* dim_bm;
  uint64_t* dbits;
  uint32_t* err;
};

struct MQStarArgs {
  const MQStarQ* q;
  uint32_t nq;
  const MQDimSlot* slots;
  // the pass's queries (or slots) by its table's row count, descending, in
  // device memory: an index into an array passed by value would make the
  // compiler hold the whole argument block in registers
  const uint8_t* ord;
  uint32_t n_ord;
  uint64_t n_max;  // that table's largest row count in the batch
  const void* d_key;
  int32_t d_key_type;
  uint32_t d_kvec;
  const void* d_pcol;
  uint32_t d_pw, d_pvec;  // predicate column width (0: none), alignment
  int64_t d_min;
  uint64_t dspan;
  const void* p_key;
  int32_t p_key_type;
  uint32_t p_kvec;
  const void* p_fk;
  int32_t p_fk_type;
  uint32_t p_fvec;
  const void* p_pcol;
  uint32_t p_pw, p_pvec;
  int64_t p_min;
  const void* f_fk;
  int32_t f_fk_type;
  uint32_t f_fvec;
  const void* f_pcol;
  uint32_t f_pw, f_pvec;
  mtl_expr expr;
};

__device__ __forceinline__ void pred_u32x16(const PredC& p, const uint32_t (&x)[FJ][4],
                                            uint32_t& bits) {
  if (p.kind == PK_I32 || p.kind == PK_U8) {  // uniform
    // an int32 value or a zero-extended byte: the compiled int32 interval
    // (three operations per row instead of the 64-bit encoding)
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (!pred_pass_i32(x[j][v], p)) bits &= ~(1u << (4 * j + v));
    return;
  }
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (!pred_pass(enc_raw(x[j][v], p.kind), p)) bits &= ~(1u << (4 * j + v));
}

// Set bit off[j][v] of bm for every row with a set bit (0xFFFFFFFF: key
// outside its range, error bit 8).  A lane's rows of a step share a word
// when the keys are sorted; the lanes of a 16-lane row whose word is the
// row's last lane's word OR their bits together first, so sorted keys cost
// one atomic per 64 keys.
__device__ __forceinline__ void mark_global(uint64_t* bm, const uint32_t (&off)[FJ][4],
                                            uint32_t bits, uint32_t* err, int lane) {
  uint32_t e = 0;
#pragma unroll
  for (int j = 0; j < FJ; ++j) {
    const uint32_t m = (bits >> (4 * j)) & 0xFu;
    uint32_t w = 0xFFFFFFFFu, rest = 0;
    uint64_t b = 0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      if (!((m >> v) & 1u)) continue;
      const uint32_t s = off[j][v];
      if (s == 0xFFFFFFFFu) {
        e |= 8u;
        continue;
      }
      if (w == 0xFFFFFFFFu) w = s >> 6;
      if ((s >> 6) == w) b |= 1ull << (s & 63);
      else rest |= 1u << v;
    }
    const uint32_t wl = (uint32_t)__shfl((int)w, lane | 15, WAVE);
    const bool join = b != 0 && w == wl;
    unsigned long long jb = join ? b : 0ull;
    jb |= __shfl_xor(jb, 1, WAVE);
    jb |= __shfl_xor(jb, 2, WAVE);
    jb |= __shfl_xor(jb, 4, WAVE);
    jb |= __shfl_xor(jb, 8, WAVE);
    if ((lane & 15) == 15 && jb) atomicOr((unsigned long long*)&bm[wl], jb);
    if (b && !join) atomicOr((unsigned long long*)&bm[w], (unsigned long long)b);
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if ((rest >> v) & 1u) {
        const uint32_t s = off[j][v];
        atomicOr((unsigned long long*)&bm[s >> 6], 1ull << (s & 63));
      }
  }
  if (e) atomicOr(err, e);
}

// The tile-major passes below take one tile per block and every query (or
// dimension slot) that covers it: the covering ones are a prefix of
// `ord`, which lists them by the pass table's row count, descending.  The
// tile's key and predicate columns are loaded once, for the table's largest
// row count in the batch, and each query then masks its own rows.

// Dimension pass: each slot's pass bits and key bitmap for one tile.
__global__ __launch_bounds__(FBLOCK) void k_mqs_dim(MQStarArgs a) {
  const uint32_t tile = blockIdx.x;
  const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  const uint64_t seg = (uint64_t)tile * FTILE + (uint64_t)wave * (FJ * 256);
  const uint64_t k0 = seg + (uint64_t)lane * 4;
  const uint64_t n = a.n_max;
  uint32_t px[FJ][4], ko[FJ][4];
  if (a.d_pw) load_u32x16(a.d_pcol, a.d_pw, k0, n, a.d_pvec, px);
  load_u32x16(a.d_key, a.d_key_type == MTL_DICT8 ? 1u : 4u, k0, n, a.d_kvec, ko);
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) ko[j][v] = narrow_off(ko[j][v], a.d_key_type, a.d_min, a.dspan);
#pragma unroll 1
  for (uint32_t i = 0; i < a.n_ord; ++i) {
    const MQDimSlot& S = a.slots[a.ord[i]];
    if ((uint64_t)tile * FTILE >= S.rows) break;  // block-uniform
    uint64_t kk = k0;
    int ln = lane;
    keep_in_loop(px);
    keep_in_loop(ko);
    keep_in_loop(kk);
    keep_in_loop(ln);
    uint32_t bits = rows_below(kk, S.rows);
    if (a.d_pw) pred_u32x16(S.pred, px, bits);
    store_pass_bits(S.dbits, seg, bits, ln);
    mark_global(S.dim_bm, ko, bits, S.err, ln);
  }
}

// Bits of the lane's rows in `live` whose dimension offset is set in bm.
__device__ __forceinline__ uint32_t probe_dim(const uint64_t* bm, const uint32_t (&fo)[FJ][4],
                                              uint32_t live) {
  const uint32_t* bm32 = (const uint32_t*)bm;
  uint32_t w[FJ][4];
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const bool need = ((live >> (4 * j + v)) & 1u) && fo[j][v] != 0xFFFFFFFFu;
      w[j][v] = need ? ld_off32(bm32, (fo[j][v] >> 5) << 2) : 0u;
    }
  uint32_t hit = 0;
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (fo[j][v] != 0xFFFFFFFFu && ((w[j][v] >> (fo[j][v] & 31)) & 1u)) hit |= 1u << (4 * j + v);
  return hit & live;
}

// Parent pass: for each query covering the tile, its predicate, the
// semi-join with its dimension slot's bitmap (probed once per slot and
// tile, on the first query of the slot), its pass bits and parent key
// bitmap.  Block 0 also carries each slot's dimension-pass errors over to
// the slot's queries.
__global__ __launch_bounds__(FBLOCK) void k_mqs_parent(MQStarArgs a) {
  const uint32_t tile = blockIdx.x;
  const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  if (tile == 0 && threadIdx.x < a.nq) {
    const MQStarQ& Q = a.q[threadIdx.x];
    const uint32_t e = *a.slots[Q.dslot].err;
    if (e) atomicOr(Q.err, e);
  }
  const uint64_t seg = (uint64_t)tile * FTILE + (uint64_t)wave * (FJ * 256);
  const uint64_t k0 = seg + (uint64_t)lane * 4;
  const uint64_t n = a.n_max;
  uint32_t px[FJ][4], fo[FJ][4], ko[FJ][4];
  if (a.p_pw) load_u32x16(a.p_pcol, a.p_pw, k0, n, a.p_pvec, px);
  load_u32x16(a.p_fk, a.p_fk_type == MTL_DICT8 ? 1u : 4u, k0, n, a.p_fvec, fo);
  load_u32x16(a.p_key, a.p_key_type == MTL_DICT8 ? 1u : 4u, k0, n, a.p_kvec, ko);
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      fo[j][v] = narrow_off(fo[j][v], a.p_fk_type, a.d_min, a.dspan);
      ko[j][v] = narrow_off(ko[j][v], a.p_key_type, a.p_min, 1ull << 32);
    }
  // the probe bits of each slot, computed on its first query in this tile
  uint32_t pb[kDimSlots];
  uint32_t have = 0;  // bit s: pb[s] is valid (uniform)
#pragma unroll
  for (uint32_t s = 0; s < kDimSlots; ++s) pb[s] = 0;
  uint32_t live = 0;
#pragma unroll 1
  for (uint32_t i = 0; i < a.n_ord; ++i) {
    const MQStarQ& Q = a.q[a.ord[i]];
    if ((uint64_t)tile * FTILE >= Q.p_rows) break;  // block-uniform
    uint64_t kk = k0;
    int ln = lane;
    keep_in_loop(px);
    keep_in_loop(fo);
    keep_in_loop(ko);
    keep_in_loop(kk);
    keep_in_loop(ln);
    if (i == 0) live = rows_below(kk, Q.p_rows);    // the largest covering query's rows
    const uint32_t sl = Q.dslot;
    if (!((have >> sl) & 1u)) {
      const uint32_t h = probe_dim(a.slots[sl].dim_bm, fo, live);
#pragma unroll
      for (uint32_t s = 0; s < kDimSlots; ++s) pb[s] = s == sl ? h : pb[s];
      have |= 1u << sl;
    }
    uint32_t bits = rows_below(kk, Q.p_rows);
    if (a.p_pw) pred_u32x16(Q.ppred, px, bits);
    uint32_t m = 0;
#pragma unroll
    for (uint32_t s = 0; s < kDimSlots; ++s) m = s == sl ? pb[s] : m;
    bits &= m;
    uint32_t kq[FJ][4];
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) kq[j][v] = ko[j][v] < Q.pspan ? ko[j][v] : 0xFFFFFFFFu;
    store_pass_bits(Q.pbits, seg, bits, ln);
    mark_global(Q.par_bm, kq, bits, Q.err, ln);
  }
}

// Fact pass: the tile's foreign keys and predicate column are loaded once;
// queries are taken two at a time (both queries' probe loads in flight
// together): each query's predicate and row count, the probe of its parent
// bitmap (one two-word window per lane and step: the facts are clustered by
// parent key, so a step's four rows name one or two parents; rows outside
// the window probe one by one), and its hits appended to the wave's LDS hit
// list.  A full list is flushed by the whole wave, one hit per lane with
// every load in flight: slot rank, expression columns, value, atomic add
// to the query's accumulator.  Overflow flags and the largest |value| per
// query gather in LDS and reach global memory once per block.
constexpr uint32_t kHitCap = 256;  // hits per wave list

struct MQHitLds {
  uint32_t key[FWAVES][kHitCap];   // query << 16 | row within the tile
  uint32_t slot[FWAVES][kHitCap];  // parent key offset
  uint64_t* acc[kStarMQMax];
  uint32_t* cnt[kStarMQMax];
  const uint64_t* par_bm[kStarMQMax];
  const uint32_t* rank[kStarMQMax];
  unsigned long long tmax[kStarMQMax];
  unsigned long long ovf;  // bit q: query q's expression overflowed
};

template <bool DBL>
__device__ __forceinline__ void mqs_flush(const MQStarArgs& a, MQHitLds& L, int wave, int lane,
                                          uint32_t n, uint64_t t0) {
  using V = typename std::conditional<DBL, double, int64_t>::type;
  for (uint32_t e0 = 0; e0 < n; e0 += WAVE) {
    const uint32_t e = e0 + (uint32_t)lane;
    if (e >= n) break;
    const uint32_t key = L.key[wave][e], sv = L.slot[wave][e];
    const uint32_t q = key >> 16;
    const uint64_t row = t0 + (key & 0xFFFFu);
    const uint64_t wd = L.par_bm[q][sv >> 6];
    const uint32_t rk = L.rank[q][sv >> 6];
    V cv[MTL_EXPR_MAX_COLS];
#pragma unroll
    for (int c = 0; c < MTL_EXPR_MAX_COLS; ++c) {
      if ((uint32_t)c < a.expr.n_cols) {
        if constexpr (DBL) cv[c] = load_f64(a.expr.col[c], a.expr.col_type[c], row);
        else cv[c] = load_key(a.expr.col[c], a.expr.col_type[c], row);
      } else {
        cv[c] = 0;
      }
    }
    const uint64_t below = (sv & 63) ? (wd & ((1ull << (sv & 63)) - 1)) : 0ull;
    const uint32_t r = rk + (uint32_t)__popcll(below);
    if constexpr (DBL) {
      unsafeAtomicAdd((double*)&L.acc[q][r], expr_f64_vals(a.expr, cv));
    } else {
      bool ovf = false;
      const int64_t val = expr_i64_vals(a.expr, cv, ovf);
      atomicAdd((unsigned long long*)&L.acc[q][r], (unsigned long long)val);
      const uint64_t av = val < 0 ? (uint64_t)0 - (uint64_t)val : (uint64_t)val;
      if (av) atomicMax(&L.tmax[q], (unsigned long long)av);
      if (ovf) atomicOr(&L.ovf, 1ull << q);
    }
    atomicAdd(&L.cnt[q][r], 1u);
  }
}

// Candidates of query Q among the lane's rows and the two-word window of
// each step (loads issued, not yet used).
struct FactProbe {
  uint32_t cand, wb[FJ], lo[FJ], hi[FJ];
};

__device__ __forceinline__ void fact_probe_issue(const MQStarArgs& a, const MQStarQ& Q,
                                                 const uint32_t (&px)[FJ][4],
                                                 const uint32_t (&fo)[FJ][4], uint64_t k0,
                                                 FactProbe& P) {
  uint32_t cand = rows_below(k0, Q.f_rows);
  if (a.f_pw) pred_u32x16(Q.fpred, px, cand);
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (fo[j][v] >= Q.pspan) cand &= ~(1u << (4 * j + v));
  P.cand = cand;
  const uint32_t* bm32 = (const uint32_t*)Q.par_bm;
#pragma unroll
  for (int j = 0; j < FJ; ++j) {
    uint32_t s0 = 0;
#pragma unroll
    for (int v = 3; v >= 0; --v) s0 = ((cand >> (4 * j + v)) & 1u) ? fo[j][v] : s0;
    P.wb[j] = s0 >> 5;
    // the parent bitmap has a zero word past its end: wb + 1 is in bounds
    P.lo[j] = ld_off32(bm32, P.wb[j] << 2);
    P.hi[j] = ld_off32(bm32, (P.wb[j] + 1) << 2);
  }
}

__device__ __forceinline__ uint32_t fact_probe_hits(const MQStarQ& Q, const FactProbe& P,
                                                    const uint32_t (&fo)[FJ][4]) {
  uint32_t hits = 0, far = 0;
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const uint32_t b = 1u << (4 * j + v);
      if (!(P.cand & b)) continue;
      const uint32_t d = fo[j][v] - (P.wb[j] << 5);
      if (d < 64) {
        if (((d < 32 ? P.lo[j] : P.hi[j]) >> (d & 31)) & 1u) hits |= b;
      } else {
        far |= b;
      }
    }
  if (__ballot(far != 0)) {  // wave-uniform
    const uint32_t* bm32 = (const uint32_t*)Q.par_bm;
    uint32_t w[FJ][4];
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        w[j][v] = ((far >> (4 * j + v)) & 1u) ? ld_off32(bm32, (fo[j][v] >> 5) << 2) : 0u;
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (((far >> (4 * j + v)) & 1u) && ((w[j][v] >> (fo[j][v] & 31)) & 1u))
          hits |= 1u << (4 * j + v);
  }
  return hits;
}

// Append query q's hits of this lane to the wave's list; *n is the list
// length (uniform).  A query adds at most 1024 hits per wave, so the list
// may fill more than once: it is filled to kHitCap, flushed, and refilled.
template <bool DBL>
__device__ __forceinline__ void mqs_append(const MQStarArgs& a, MQHitLds& L, int wave, int lane,
                                           uint32_t& n, uint32_t q, uint32_t hits,
                                           const uint32_t (&fo)[FJ][4], uint64_t t0,
                                           uint32_t seg_row) {
  const uint32_t c = (uint32_t)__popc(hits);
  const uint32_t incl = wave_inclusive_u32(c);
  uint32_t end = n + (uint32_t)__builtin_amdgcn_readlane((int)incl, WAVE - 1);  // uniform
  uint32_t p = n + incl - c;  // list position of the lane's next hit
  uint32_t h = hits;
  while (true) {
    while (h && p < kHitCap) {
      const int i = __ffs(h) - 1;
      h &= h - 1;
      uint32_t sv = fo[0][0];
#pragma unroll
      for (int t = 1; t < FJ * 4; ++t) sv = i == t ? fo[t >> 2][t & 3] : sv;
      L.key[wave][p] = (q << 16) | (seg_row + (uint32_t)(i >> 2) * 256 + (uint32_t)(i & 3));
      L.slot[wave][p] = sv;
      ++p;
    }
    if (end <= kHitCap) break;
    // full: entries written by other lanes of this wave are read below
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    mqs_flush<DBL>(a, L, wave, lane, kHitCap, t0);
    __builtin_amdgcn_wave_barrier();
    p -= kHitCap;  // lanes with hits left were at or past the end
    end -= kHitCap;
  }
  n = end;
}

template <bool DBL>
__global__ __launch_bounds__(FBLOCK) void k_mqs_fact(MQStarArgs a) {
  __shared__ MQHitLds L;
  const uint32_t tile = blockIdx.x;
  const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  if (threadIdx.x < a.nq) {
    const MQStarQ& Q = a.q[threadIdx.x];
    L.acc[threadIdx.x] = Q.acc;
    L.cnt[threadIdx.x] = Q.cnt;
    L.par_bm[threadIdx.x] = Q.par_bm;
    L.rank[threadIdx.x] = Q.rank;
    L.tmax[threadIdx.x] = 0;
  }
  if (threadIdx.x == 0) L.ovf = 0;
  __syncthreads();
  const uint64_t t0 = (uint64_t)tile * FTILE;
  const uint32_t seg_row = (uint32_t)wave * (FJ * 256) + (uint32_t)lane * 4;
  const uint64_t k0 = t0 + seg_row;
  const uint64_t nmax = a.n_max;
  uint32_t px[FJ][4], fo[FJ][4];
  if (a.f_pw) load_u32x16(a.f_pcol, a.f_pw, k0, nmax, a.f_pvec, px);
  load_u32x16(a.f_fk, a.f_fk_type == MTL_DICT8 ? 1u : 4u, k0, nmax, a.f_fvec, fo);
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) fo[j][v] = narrow_off(fo[j][v], a.f_fk_type, a.p_min, 1ull << 32);
  uint32_t n = 0;  // the wave's list length
#pragma unroll 1
  for (uint32_t i = 0; i < a.n_ord; ++i) {
    const uint32_t q = a.ord[i];
    const MQStarQ& Q = a.q[q];
    if (t0 >= Q.f_rows) break;  // block-uniform
    uint64_t kk = k0;
    int ln = lane;
    uint32_t sr = seg_row;
    keep_in_loop(px);
    keep_in_loop(fo);
    keep_in_loop(kk);
    keep_in_loop(ln);
    keep_in_loop(sr);
    FactProbe P;
    fact_probe_issue(a, Q, px, fo, kk, P);
    const uint32_t h = fact_probe_hits(Q, P, fo);
    mqs_append<DBL>(a, L, wave, ln, n, q, h, fo, t0, sr);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (n) mqs_flush<DBL>(a, L, wave, lane, n, t0);
  __syncthreads();
  if (threadIdx.x < a.nq) {
    const MQStarQ& Q = a.q[threadIdx.x];
    if (L.tmax[threadIdx.x]) atomicMax(Q.tmax, L.tmax[threadIdx.x]);
    if ((L.ovf >> threadIdx.x) & 1ull) atomicOr(Q.err, 16u);
  }
}

static bool narrow_key_type(int32_t t) {
  return t == MTL_INTEGER || t == MTL_DATE || t == MTL_DICT32 || t == MTL_DICT8;
}

static bool narrow_pred(const mtl_pred* p, uint32_t n) {
  return n == 0 || (n == 1 && type_width(p[0].type) >= 1 && type_width(p[0].type) <= 4);
}

bool star_batchable(const mtl_star_desc& d) {
  StarPlan sp;
  if (star_validate(&d, &sp) != MTL_OK) return false;
  return d.d_rows && narrow_key_type(d.d_key_type) && narrow_key_type(d.p_key_type) &&
         narrow_key_type(d.p_fk_type) && narrow_key_type(d.f_fk_type) &&
         narrow_pred(d.d_preds, d.d_n_preds) && narrow_pred(d.p_preds, d.p_n_preds) &&
         narrow_pred(d.f_preds, d.f_n_preds);
}

static bool same_pred_col(const mtl_pred* a, uint32_t na, const mtl_pred* b, uint32_t nb) {
  return na == nb && (na == 0 || (a[0].column == b[0].column && a[0].type == b[0].type));
}

bool star_shareable(const mtl_star_desc& a, const mtl_star_desc& b) {
  return a.d_key == b.d_key && a.d_key_type == b.d_key_type && a.d_key_min == b.d_key_min &&
         a.d_key_max == b.d_key_max && a.p_key == b.p_key && a.p_key_type == b.p_key_type &&
         a.p_key_min == b.p_key_min && a.p_fk == b.p_fk && a.p_fk_type == b.p_fk_type &&
         a.f_fk == b.f_fk && a.f_fk_type == b.f_fk_type &&
         same_pred_col(a.d_preds, a.d_n_preds, b.d_preds, b.d_n_preds) &&
         same_pred_col(a.p_preds, a.p_n_preds, b.p_preds, b.p_n_preds) &&
         same_pred_col(a.f_preds, a.f_n_preds, b.f_preds, b.f_n_preds) &&
         std::memcmp(&a.agg, &b.agg, sizeof(mtl_expr)) == 0;
}

bool star_dim_same(const mtl_star_desc& a, const mtl_star_desc& b) {
  if (a.d_rows != b.d_rows || a.d_n_preds != b.d_n_preds) return false;
  return a.d_n_preds == 0 ||
         (a.d_preds[0].op == b.d_preds[0].op && a.d_preds[0].scalar == b.d_preds[0].scalar);
}

// Merge levels of one query's top-K: lists per level, from its phase-1 grid.
static uint32_t merge_levels(uint32_t grid, uint32_t k) {
  const uint32_t fan = topk_fan(k);
  uint32_t L = grid, lv = 0;
  while (true) {
    ++lv;
    const uint32_t blocks = (uint32_t)ceil_div(L, fan);
    if (blocks == 1) return lv;
    L = blocks;
  }
}

size_t star_batch_stage_bytes(uint32_t n) {
  // per query: its pass arguments, rank, map and top-K arguments, and up to
  // 8 merge levels
  return (size_t)n * (sizeof(MQStarQ) + sizeof(RankLaunch) + sizeof(MapArgs) + sizeof(TopkArgs) +
                      8 * sizeof(MQMergeQ)) +
         kDimSlots * sizeof(MQDimSlot) + 1024;
}

int32_t star_batch_enqueue(const StarBatchItem* items, uint32_t n, void* host_stage,
                           mtl_stream* s) {
  if (n == 0) return MTL_OK;
  if (n > kStarMQMax) return set_error(MTL_ERR_INVALID_ARGUMENT, "more than %u queries", kStarMQMax);
  const mtl_star_desc& d0 = *items[0].d;
  std::vector<StarPlan> sp(n);
  uint32_t levels = 0;
  for (uint32_t i = 0; i < n; ++i) {
    MTL_TRY(star_validate(items[i].d, &sp[i]));
    if (!star_batchable(*items[i].d) || !star_shareable(d0, *items[i].d))
      return set_error(MTL_ERR_INVALID_ARGUMENT, "query %u cannot join the shared passes", i);
    const uint32_t lv = merge_levels(sp[i].topk_grid, items[i].d->k);
    if (lv > 8) return set_error(MTL_UNSUPPORTED, "top-K of %u merge levels", lv);
    levels += lv;
  }
  // dimension slots: one per distinct dimension predicate (and row count)
  std::vector<uint32_t> slot_of(n), slot_rep;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t sl = 0;
    while (sl < slot_rep.size() && !star_dim_same(*items[slot_rep[sl]].d, *items[i].d)) ++sl;
    if (sl == slot_rep.size()) {
      if (sl == kDimSlots)
        return set_error(MTL_ERR_INVALID_ARGUMENT, "more than %u dimension predicates",
                         kDimSlots);
      slot_rep.push_back(i);
    }
    slot_of[i] = sl;
  }
  const uint32_t n_slots = (uint32_t)slot_rep.size();
  Arena& ar = s->arena;
  hipStream_t st = s->stream;
  // regions to clear, carved from one block (each a multiple of 64 words)
  auto pad = [](uint64_t w) { return (w + 63) & ~uint64_t(63); };
  std::vector<StarBufs> b(n);
  uint64_t clear_words = 0;
  for (uint32_t sl = 0; sl < n_slots; ++sl) clear_words += pad(sp[slot_rep[sl]].dwords * 2) + 64;
  for (uint32_t i = 0; i < n; ++i) {
    b[i].rtiles = (uint32_t)ceil_div(sp[i].pwords, RTILE);
    clear_words += pad((sp[i].pwords + 1) * 2) + 64 +
                   pad((uint64_t)std::max<uint32_t>(b[i].rtiles, 1) * 2) + 64;
  }
  void* clr;
  MTL_TRY(ar.alloc(clear_words * 4, &clr));
  uint32_t* cw = (uint32_t*)clr;
  // the slots' bitmaps (cleared) and pass bits
  std::vector<MQDimSlot> slots(n_slots);
  for (uint32_t sl = 0; sl < n_slots; ++sl) {
    const mtl_star_desc& d = *items[slot_rep[sl]].d;
    MQDimSlot& S = slots[sl];
    S.pred = d.d_n_preds ? compile_pred(d.d_preds[0]) : PredC{};
    S.rows = d.d_rows;
    S.dim_bm = (uint64_t*)cw;
    cw += pad(sp[slot_rep[sl]].dwords * 2);
    S.err = cw;
    cw += 64;
    void* db;
    MTL_TRY(ar.alloc((uint64_t)filter_tiles(d.d_rows) * kTileRows / 8, &db));
    S.dbits = (uint64_t*)db;
  }
  // staged arguments: [MQStarQ n][RankLaunch n][MapArgs n][TopkArgs n, by
  // top-K shape][MQMergeQ, by level][MQDimSlot n_slots]
  const size_t o_r = (size_t)n * sizeof(MQStarQ);
  const size_t o_m = o_r + (size_t)n * sizeof(RankLaunch);
  const size_t o_t = o_m + (size_t)n * sizeof(MapArgs);
  const size_t o_g = o_t + (size_t)n * sizeof(TopkArgs);
  const size_t o_s = o_g + (size_t)levels * sizeof(MQMergeQ);
  const size_t o_o = o_s + (size_t)n_slots * sizeof(MQDimSlot);  // 3 order lists
  const size_t stage_bytes = o_o + 3 * (size_t)kStarMQMax;
  void* dq;
  MTL_TRY(ar.alloc(stage_bytes, &dq));
  char* hs = (char*)host_stage;
  MQStarQ* hq = (MQStarQ*)hs;
  RankLaunch* hr = (RankLaunch*)(hs + o_r);
  MapArgs* hm = (MapArgs*)(hs + o_m);
  TopkArgs* ht = (TopkArgs*)(hs + o_t);
  MQMergeQ* hg = (MQMergeQ*)(hs + o_g);
  std::vector<TopkArgs> ta(n);
  uint32_t rgrid = 1, mgrid = 1;
  for (uint32_t i = 0; i < n; ++i) {
    const mtl_star_desc& d = *items[i].d;
    StarBufs& bi = b[i];
    bi.dim_bm = slots[slot_of[i]].dim_bm;
    bi.dbits = slots[slot_of[i]].dbits;
    bi.par_bm = cw;
    cw += pad((sp[i].pwords + 1) * 2);
    bi.misc = cw;
    cw += 64;
    bi.st_r = cw;
    cw += pad((uint64_t)std::max<uint32_t>(bi.rtiles, 1) * 2);
    bi.tmax = cw;
    cw += 64;
    bi.n_tmax = 1;
    const uint64_t prows = d.p_rows ? d.p_rows : 1;
    MTL_TRY(ar.alloc(sp[i].pwords * 4, &bi.par_rank));
    MTL_TRY(ar.alloc((uint64_t)filter_tiles(d.p_rows) * kTileRows / 8, &bi.pbits));
    MTL_TRY(ar.alloc(prows * 4, &bi.row_of));
    MTL_TRY(ar.alloc(prows * 8, &bi.acc));
    MTL_TRY(ar.alloc(prows * 4, &bi.cnt));
    MTL_TRY(alloc_topk_bufs(&d, sp[i], ar, bi));
    MQStarQ& Q = hq[i];
    Q = MQStarQ{};
    Q.d_rows = d.d_rows;
    Q.p_rows = d.p_rows;
    Q.f_rows = d.f_rows;
    if (d.d_n_preds) Q.dpred = compile_pred(d.d_preds[0]);
    if (d.p_n_preds) Q.ppred = compile_pred(d.p_preds[0]);
    if (d.f_n_preds) Q.fpred = compile_pred(d.f_preds[0]);
    Q.pspan = sp[i].pspan;
    Q.dslot = slot_of[i];
    Q.dim_bm = (uint64_t*)bi.dim_bm;
    Q.dbits = (uint64_t*)bi.dbits;
    Q.pbits = (uint64_t*)bi.pbits;
    Q.par_bm = (uint64_t*)bi.par_bm;
    Q.rank = (const uint32_t*)bi.par_rank;
    Q.acc = (uint64_t*)bi.acc;
    Q.cnt = (uint32_t*)bi.cnt;
    Q.err = (uint32_t*)bi.misc + 3;
    Q.tmax = (unsigned long long*)bi.tmax;
    hr[i] = make_rank(&d, sp[i], bi);
    hm[i] = make_map(&d, sp[i], bi);
    ta[i] = make_topk(&d, sp[i], bi, items[i].out, items[i].out_n);
    rgrid = std::max(rgrid, hr[i].grid);
    mgrid = std::max(mgrid, hm[i].grid);
  }
  // top-K arguments grouped by wave-list shape (K <= 64, <= 128, above)
  auto shape = [](uint32_t k) { return k <= 64 ? 0 : k <= 128 ? 1 : 2; };
  std::vector<uint32_t> tpos(n);  // query -> position in ht
  uint32_t cls_off[4] = {0, 0, 0, 0}, cls_grid[3] = {1, 1, 1};
  {
    uint32_t at = 0;
    for (int c = 0; c < 3; ++c) {
      cls_off[c] = at;
      for (uint32_t i = 0; i < n; ++i)
        if (shape(items[i].d->k) == c) {
          tpos[i] = at;
          ht[at++] = ta[i];
          cls_grid[c] = std::max(cls_grid[c], ta[i].grid);
        }
    }
    cls_off[3] = at;
  }
  // merge levels, level-major: every query's level j before any level j+1
  std::vector<uint32_t> lvl_off, lvl_grid;
  {
    std::vector<uint32_t> L(n), in(n, 0), left(n);
    for (uint32_t i = 0; i < n; ++i) {
      L[i] = sp[i].topk_grid;
      left[i] = merge_levels(sp[i].topk_grid, items[i].d->k);
    }
    uint32_t at = 0;
    while (true) {
      const uint32_t start = at;
      uint32_t g = 1;
      for (uint32_t i = 0; i < n; ++i) {
        if (!left[i]) continue;
        const uint32_t fan = topk_fan(items[i].d->k);
        const uint32_t blocks = (uint32_t)ceil_div(L[i], fan);
        MQMergeQ& m = hg[at++];
        m.ta = (const TopkArgs*)((char*)dq + o_t) + tpos[i];
        m.in = TopkLists{(uint64_t*)b[i].lk0[in[i]], (uint32_t*)b[i].lg[in[i]],
                         (uint32_t*)b[i].ln[in[i]]};
        m.out = TopkLists{(uint64_t*)b[i].lk0[in[i] ^ 1], (uint32_t*)b[i].lg[in[i] ^ 1],
                          (uint32_t*)b[i].ln[in[i] ^ 1]};
        m.L = L[i];
        m.final_level = blocks == 1 ? 1u : 0u;
        m.blocks = blocks;
        g = std::max(g, blocks);
        L[i] = blocks;
        in[i] ^= 1;
        --left[i];
      }
      if (at == start) break;
      lvl_off.push_back(start);
      lvl_off.push_back(at);
      lvl_grid.push_back(g);
    }
  }
  std::memcpy(hs + o_s, slots.data(), (size_t)n_slots * sizeof(MQDimSlot));
  // each pass's queries (slots) by its table's row count, descending
  auto order = [&](int pass, auto rows_of, uint32_t cnt) {
    std::vector<uint32_t> ix(cnt);
    for (uint32_t i = 0; i < cnt; ++i) ix[i] = i;
    std::stable_sort(ix.begin(), ix.end(),
                     [&](uint32_t x, uint32_t y) { return rows_of(x) > rows_of(y); });
    uint8_t* o = (uint8_t*)hs + o_o + (size_t)pass * kStarMQMax;
    for (uint32_t i = 0; i < cnt; ++i) o[i] = (uint8_t)ix[i];
    return ix[0];
  };
  const uint32_t dim_first = order(0, [&](uint32_t sl) { return slots[sl].rows; }, n_slots);
  const uint32_t par_first = order(1, [&](uint32_t i) { return items[i].d->p_rows; }, n);
  const uint32_t fact_first = order(2, [&](uint32_t i) { return items[i].d->f_rows; }, n);
  MTL_HIP(hipMemcpyAsync(dq, hs, stage_bytes, hipMemcpyHostToDevice, st));
  ZeroList zl;
  zl.add((uint32_t*)clr, clear_words);
  MTL_TRY(launch_zero(zl, st));
  auto vec = [](const void* p, uint32_t w) {
    return ((uintptr_t)p % (w == 1 ? 4 : 16)) == 0 ? 1u : 0u;
  };
  auto kw = [](int32_t t) { return t == MTL_DICT8 ? 1u : 4u; };
  MQStarArgs a{};
  a.q = (const MQStarQ*)dq;
  a.d_key = d0.d_key;
  a.d_key_type = d0.d_key_type;
  a.d_kvec = vec(d0.d_key, kw(d0.d_key_type));
  a.d_pw = d0.d_n_preds ? (uint32_t)type_width(d0.d_preds[0].type) : 0u;
  a.d_pcol = d0.d_n_preds ? d0.d_preds[0].column : nullptr;
  a.d_pvec = a.d_pw ? vec(a.d_pcol, a.d_pw) : 0u;
  a.d_min = d0.d_key_min;
  a.dspan = sp[0].dspan;
  a.p_key = d0.p_key;
  a.p_key_type = d0.p_key_type;
  a.p_kvec = vec(d0.p_key, kw(d0.p_key_type));
  a.p_fk = d0.p_fk;
  a.p_fk_type = d0.p_fk_type;
  a.p_fvec = vec(d0.p_fk, kw(d0.p_fk_type));
  a.p_pw = d0.p_n_preds ? (uint32_t)type_width(d0.p_preds[0].type) : 0u;
  a.p_pcol = d0.p_n_preds ? d0.p_preds[0].column : nullptr;
  a.p_pvec = a.p_pw ? vec(a.p_pcol, a.p_pw) : 0u;
  a.p_min = d0.p_key_min;
  a.f_fk = d0.f_fk;
  a.f_fk_type = d0.f_fk_type;
  a.f_fvec = vec(d0.f_fk, kw(d0.f_fk_type));
  a.f_pw = d0.f_n_preds ? (uint32_t)type_width(d0.f_preds[0].type) : 0u;
  a.f_pcol = d0.f_n_preds ? d0.f_preds[0].column : nullptr;
  a.f_pvec = a.f_pw ? vec(a.f_pcol, a.f_pw) : 0u;
  a.expr = d0.agg;
  a.nq = n;
  a.slots = (const MQDimSlot*)((char*)dq + o_s);
  {
    ProfScope ps(s, "k_mqs_dim");
    a.ord = (const uint8_t*)dq + o_o;
    a.n_ord = n_slots;
    a.n_max = slots[dim_first].rows;
    hipLaunchKernelGGL(k_mqs_dim, dim3(filter_tiles(a.n_max)), dim3(FBLOCK), 0, st, a);
    MTL_HIP(hipGetLastError());
  }
  {
    ProfScope ps(s, "k_mqs_parent");
    a.ord = (const uint8_t*)dq + o_o + kStarMQMax;
    a.n_ord = n;
    a.n_max = items[par_first].d->p_rows;
    hipLaunchKernelGGL(k_mqs_parent, dim3(std::max(1u, filter_tiles(a.n_max))), dim3(FBLOCK), 0,
                       st, a);
    MTL_HIP(hipGetLastError());
  }
  {
    ProfScope ps(s, "k_mq_rank");
    hipLaunchKernelGGL(k_mq_rank, dim3(rgrid, n), dim3(RBLOCK), 0, st,
                       (const RankLaunch*)((char*)dq + o_r));
    MTL_HIP(hipGetLastError());
  }
  {
    ProfScope ps(s, "k_mq_map");
    const MapArgs* dm = (const MapArgs*)((char*)dq + o_m);
    if (map_k4(&d0)) hipLaunchKernelGGL(k_mq_map<true>, dim3(mgrid, n), dim3(256), 0, st, dm);
    else hipLaunchKernelGGL(k_mq_map<false>, dim3(mgrid, n), dim3(256), 0, st, dm);
    MTL_HIP(hipGetLastError());
  }
  a.ord = (const uint8_t*)dq + o_o + 2 * kStarMQMax;
  a.n_ord = n;
  a.n_max = items[fact_first].d->f_rows;
  if (a.n_max) {
    ProfScope ps(s, "k_mqs_fact");
    const dim3 grid(filter_tiles(a.n_max));
    if (d0.agg.is_double) hipLaunchKernelGGL(k_mqs_fact<true>, grid, dim3(FBLOCK), 0, st, a);
    else hipLaunchKernelGGL(k_mqs_fact<false>, grid, dim3(FBLOCK), 0, st, a);
    MTL_HIP(hipGetLastError());
  }
  {
    ProfScope ps(s, "k_mq_star_topk");
    const TopkArgs* dt_ = (const TopkArgs*)((char*)dq + o_t);
    for (int c = 0; c < 3; ++c) {
      const uint32_t cnt = cls_off[c + 1] - cls_off[c];
      if (!cnt) continue;
      const dim3 g(cls_grid[c], cnt);
      if (c == 0)
        hipLaunchKernelGGL(k_mq_star_topk<1>, g, dim3(TopkShape<1>::NT), 0, st, dt_ + cls_off[c]);
      else if (c == 1)
        hipLaunchKernelGGL(k_mq_star_topk<2>, g, dim3(TopkShape<2>::NT), 0, st, dt_ + cls_off[c]);
      else
        hipLaunchKernelGGL(k_mq_star_topk<8>, g, dim3(TopkShape<8>::NT), 0, st, dt_ + cls_off[c]);
      MTL_HIP(hipGetLastError());
    }
  }
  {
    ProfScope ps(s, "k_mq_star_merge");
    const MQMergeQ* dg = (const MQMergeQ*)((char*)dq + o_g);
    for (size_t j = 0; j < lvl_grid.size(); ++j) {
      const uint32_t lo = lvl_off[2 * j], hi = lvl_off[2 * j + 1];
      hipLaunchKernelGGL(k_mq_star_merge, dim3(lvl_grid[j], hi - lo), dim3(TOPK_MERGE_THREADS),
                         0, st, dg + lo);
      MTL_HIP(hipGetLastError());
    }
  }
  return MTL_OK;
}

}  // namespace mtl
