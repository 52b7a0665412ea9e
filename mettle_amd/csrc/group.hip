// Grouped aggregate engine: GROUP BY integer keys -> COUNT / SUM / MIN /
// MAX / AVG, plus top-K over the groups.
//
// Reference: grouped_aggregate_build_keys_v and grouped_aggregate_update_*_v
// (kernels.metal:417-621) with helpers grouped_key_hash / probe_grouped_slot
// (:291-350); host grouped_aggregate_exact_gv (aggregate_ops.rs:101-457);
// top-K topk_dense_grouped_rows{,_parallel_v} (:1162-1390).  The reference
// did one pass per aggregate, float32 CAS-loop atomics, a slot state 1
// (busy) that other threads spin on, and a host scan of the whole table.
// Here:
//   - keys are packed by their [min, max] ranges into one 64-bit value
//     (first key most significant, so packed order = key order); a packed
//     domain up to kDenseMax is indexed directly, larger domains use an
//     open-addressing table whose key is published by one 64-bit CAS (no
//     busy state, so no lane ever waits on another lane of its wave);
//   - every block keeps an LDS cache of hot keys with their partial states
//     (Zipf heads collapse there) and flushes it to the global table once;
//   - one pass updates all aggregates of a row (state words of a group are
//     contiguous, one cache line for typical aggregate lists);
//   - integer SUM is exact 128-bit with non-returning atomics only: each
//     value is split as v = lo + hi * 2^32 with lo its sign-extended low
//     half, the lo parts are summed in one signed 64-bit word and the hi
//     parts in another; sum = high * 2^32 + low, exact for fewer than 2^32
//     rows in any order, and values within int32 (hi = 0) skip the high
//     add; FLOAT/DOUBLE sums accumulate in double;
//   - in LDS, MIN/MAX read the partial state first and issue the atomic
//     only when the row improves it (rarely, after a group's first rows);
//   - each thread loads the keys and values of several rows before its
//     first update, so the loads overlap instead of waiting behind atomics.
#include <algorithm>
#include <cmath>

#include <cstring>
#include <vector>

#include "device.hpp"
#include "filter.hpp"
#include "mtl_internal.hpp"
#include "prof.hpp"
#include "topk.hpp"

namespace mtl {
using namespace dev;

constexpr uint64_t GEMPTY = ~0ull;
constexpr uint64_t kDenseMax = 1ull << 26;  // directly indexed packed domain
constexpr int kMaxW = 1 + 2 * MTL_GROUP_MAX_AGGS;
constexpr int kMaxPay = 2 * MTL_GROUP_MAX_AGGS;  // payload words of a spilled row

enum Kind : uint32_t { K_COUNT = 0, K_SUMI, K_SUMF, K_MINI, K_MAXI, K_MINF, K_MAXF };

struct AggSlot {
  uint32_t kind;
  uint32_t word;
  int32_t type;
  uint32_t dup;  // shares the state words of an earlier slot (same kind and column)
  const void* col;
  uint32_t src;  // value source (see GroupArgs::src_*)
};

// How a value source reads a column: int64, double bits, or the
// order-preserving encoding of the double (float MIN/MAX).
enum SrcConv : uint32_t { SC_INT = 0, SC_DBL, SC_ENCF };

struct GroupArgs {
  uint64_t n_rows;
  const uint32_t* sel;
  const uint32_t* sel_count;
  uint64_t sel_cap;
  uint32_t n_keys;
  uint32_t raw;  // single key used unpacked (hash table only)
  const void* key_col[MTL_GROUP_MAX_KEYS];
  int32_t key_type[MTL_GROUP_MAX_KEYS];
  int64_t key_min[MTL_GROUP_MAX_KEYS];
  uint64_t key_span[MTL_GROUP_MAX_KEYS];
  uint64_t key_stride[MTL_GROUP_MAX_KEYS];
  uint32_t n_aggs;
  AggSlot aggs[MTL_GROUP_MAX_AGGS];
  // the state updates of a row besides its count: one per aggregate that
  // owns state words (not COUNT, not a duplicate), packed as kind | word << 4
  // | source << 12
  uint32_t n_ops;
  uint32_t ops[MTL_GROUP_MAX_AGGS];  // ordered by source
  uint32_t op_beg[MTL_GROUP_MAX_AGGS + 1];  // ops of source j: [op_beg[j], op_beg[j + 1])
  // distinct (column, conversion) value sources of the aggregates: each is
  // loaded once per row (SUM, MIN and MAX of one column share one load)
  uint32_t n_src;
  const void* src_col[MTL_GROUP_MAX_AGGS];
  int32_t src_type[MTL_GROUP_MAX_AGGS];
  uint32_t src_conv[MTL_GROUP_MAX_AGGS];
  // byte width of each source's values and of the first key's; sources past
  // n_src read the first key column (always present) so every load of a
  // batch can be issued without a guard
  uint32_t src_w[MTL_GROUP_MAX_AGGS];
  uint32_t key_w0;
  uint32_t W;
  uint32_t dense;
  uint64_t cap;     // dense: domain size; hash: power-of-two capacity
  uint64_t* tkeys;  // hash: cap + 1 keys ([cap] = side slot for GEMPTY)
  uint64_t* st;     // (cap + 1) * W state words
  uint64_t init[kMaxW];
  uint32_t cache_slots;
  uint32_t cache_direct;  // dense domain that fits the cache: slot = key
  uint32_t door;          // hashed cache admits a group on its second sighting
  uint32_t* err;    // 64: table full, 8: key outside its stated range
  // spill mode (large directly indexed domains): rows whose group misses
  // the block's LDS cache are written to the block's region of the spill
  // arrays (region of block b starts at row b * per) instead of updating
  // the global table with device-scope atomics
  uint32_t spill;
  // a spilled row is its packed key (spill_key[o]) and a payload record of
  // n_pay 32-bit words (spill_pay[o * n_pay ...]): each source's value in
  // one word (a source of <= 4 bytes, narrowed exactly, see narrow_src) or
  // two (its low and high halves, words pay_lo[j] and pay_hi[j]); pay_hi ==
  // pay_lo marks a narrowed source, and sources past n_src name word 0
  uint32_t n_pay;
  uint32_t pay_lo[MTL_GROUP_MAX_AGGS];
  uint32_t pay_hi[MTL_GROUP_MAX_AGGS];
  uint32_t* spill_key;  // [rows]
  uint32_t* spill_pay;  // [rows * n_pay]
  uint32_t* spill_n;     // [grid + 1]: spilled rows of each block, then the
                         // region length (rows per block)
  uint64_t spill_rows;   // row capacity of the spill arrays
  // spill mode: each block's LDS cache is staged here at the end (slot i of
  // block b at b * cache_slots + i: its key, GEMPTY when unused, and its W
  // state words) and merged into the table after the spilled rows are
  // aggregated, since that aggregation writes whole table entries
  uint64_t* hot_key;
  uint64_t* hot_st;
  // directly indexed cache of a large input: each block's cache states are
  // staged here at the end (word k of slot i of block b at
  // (b * W + k) * (cache_slots + 1) + i) and reduced by k_dense_reduce,
  // instead of every block merging its whole cache into the table with
  // device-scope atomics
  uint64_t* dstage;
};

__device__ __forceinline__ uint64_t gfmix(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

// The exact 128-bit integer sum held as two signed 64-bit sums X (of the
// sign-extended low halves) and Y (of the high parts): value = X + Y * 2^32,
// returned as (low 64 bits, high 64 bits) of the 128-bit value.
__device__ __forceinline__ void sum128(const uint64_t* w, uint64_t& lo, int64_t& hi) {
  const int64_t X = (int64_t)w[0], Y = (int64_t)w[1];
  const uint64_t yl = (uint64_t)Y << 32;
  lo = yl + (uint64_t)X;
  hi = (Y >> 32) + (X < 0 ? -1 : 0) + (lo < yl ? 1 : 0);
}

// v = lo + hi * 2^32 with lo = (int32)v: |lo|, |hi| <= 2^31.
__device__ __forceinline__ void split_i64(int64_t v, int64_t& lo, int64_t& hi) {
  lo = (int64_t)(int32_t)(uint32_t)(uint64_t)v;
  hi = (v >> 32) + (lo < 0 ? 1 : 0);
}

// Raw bits of GE rows of a column of `width`-byte values.  The width switch
// sits outside the row loop and nothing converts the loaded registers here,
// so the GE loads go out back to back (a conversion right after each load
// would make the compiler wait for it before issuing the next).
template <int GE_>
__device__ __forceinline__ void load_bits(const void* col, uint32_t width,
                                          const uint64_t (&row)[GE_], uint64_t (&b)[GE_]) {
  if (width == 8) {
#pragma unroll
    for (int e = 0; e < GE_; ++e) b[e] = ((const uint64_t*)col)[row[e]];
  } else if (width == 4) {
#pragma unroll
    for (int e = 0; e < GE_; ++e) b[e] = ((const uint32_t*)col)[row[e]];
  } else {
#pragma unroll
    for (int e = 0; e < GE_; ++e) b[e] = ((const uint8_t*)col)[row[e]];
  }
}

// Register type holding a raw value of a WIDTH-byte column in flight (a
// 4-byte value is widened only when used, so no copy waits for its load).
template <int WIDTH> struct RawT { using type = uint64_t; };
template <> struct RawT<4> { using type = uint32_t; };
template <int GE_, int WIDTH>
__device__ __forceinline__ void load_raw(const void* col, const uint64_t (&row)[GE_],
                                         typename RawT<WIDTH>::type (&b)[GE_]) {
  using T = typename RawT<WIDTH>::type;
#pragma unroll
  for (int e = 0; e < GE_; ++e) b[e] = ((const T*)col)[row[e]];
}

// The int64 value of the raw bits of an integer column of type t.
__device__ __forceinline__ int64_t bits_key(uint64_t b, int32_t t) {
  switch (t) {
    case MTL_BIGINT:
    case MTL_TIMESTAMP:
    case MTL_DECIMAL:
      return (int64_t)b;
    case MTL_DICT8:
      return (int64_t)(b & 0xFFu);
    default:
      return (int64_t)(int32_t)(uint32_t)b;
  }
}

// A double from the raw bits of a column of type t (FLOAT, DOUBLE or integer).
__device__ __forceinline__ double bits_f64(uint64_t b, int32_t t) {
  switch (t) {
    case MTL_FLOAT: return (double)__uint_as_float((uint32_t)b);
    case MTL_DOUBLE: return __longlong_as_double((long long)b);
    default: return (double)bits_key(b, t);
  }
}

// The value source j gives for a row, from its raw bits: int64 bits
// (integer aggregates), double bits (K_SUMF), or the order-preserving
// encoding (K_MINF/K_MAXF).
__device__ __forceinline__ uint64_t src_bits(const GroupArgs& a, uint32_t j, uint64_t b) {
  switch (a.src_conv[j]) {
    case SC_INT: return (uint64_t)bits_key(b, a.src_type[j]);
    case SC_DBL: return (uint64_t)__double_as_longlong(bits_f64(b, a.src_type[j]));
    default: return enc_f64(bits_f64(b, a.src_type[j]));
  }
}

// A converted source value of a column of <= 4 bytes in 32 bits, exactly:
// an integer value fits int32, a FLOAT's double (or its order-preserving
// encoding) is a float.  widen_src inverts it.
__device__ __forceinline__ uint32_t narrow_src(const GroupArgs& a, uint32_t j, uint64_t x) {
  switch (a.src_conv[j]) {
    case SC_INT: return (uint32_t)x;
    case SC_DBL: return __float_as_uint((float)__longlong_as_double((long long)x));
    default: {  // SC_ENCF: decode the order-preserving encoding first
      const uint64_t b = (x & 0x8000000000000000ull) ? (x & ~0x8000000000000000ull) : ~x;
      return __float_as_uint((float)__longlong_as_double((long long)b));
    }
  }
}
__device__ __forceinline__ uint64_t widen_src(const GroupArgs& a, uint32_t j, uint32_t u) {
  switch (a.src_conv[j]) {
    case SC_INT: return (uint64_t)(int64_t)(int32_t)u;
    case SC_DBL: return (uint64_t)__double_as_longlong((double)__uint_as_float(u));
    default: return enc_f64((double)__uint_as_float(u));
  }
}

// The group key of a row: the key columns packed by their [min, max] ranges
// (a single raw key is used as is); false when a key lies outside its range.
// The first key column's raw bits come loaded; other key columns, if any,
// are loaded here.
__device__ __forceinline__ bool pack_key_bits(const GroupArgs& a, uint64_t row, uint64_t b0,
                                              uint64_t& key) {
  if (a.raw) {
    key = (uint64_t)bits_key(b0, a.key_type[0]);
    return true;
  }
  uint64_t off = key_off(bits_key(b0, a.key_type[0]), a.key_min[0]);
  if (off >= a.key_span[0]) return false;
  uint64_t p = off * a.key_stride[0];
  for (uint32_t i = 1; i < a.n_keys; ++i) {
    off = key_off(load_key(a.key_col[i], a.key_type[i], row), a.key_min[i]);
    if (off >= a.key_span[i]) return false;
    p += off * a.key_stride[i];
  }
  key = p;
  return true;
}

// v[s] for a wave-uniform source index s, without indexing registers
// dynamically.
template <int NS>
__device__ __forceinline__ uint64_t pick_src(const uint64_t (&v)[NS], uint32_t s) {
  uint64_t r = 0;
#pragma unroll
  for (int j = 0; j < NS; ++j) r |= v[j] & (0ull - (uint64_t)(s == (uint32_t)j));
  return r;
}

// A possibly stale read of an LDS state word.  A relaxed workgroup-scope
// atomic load becomes an LDS read; a volatile read would stay a generic
// (flat) load, which also waits for every global load in flight.
__device__ __forceinline__ uint64_t lds_peek(uint64_t* w) {
  return __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Per-row update of one group's state words (LDS or global: atomics with
// the address space of p; none returns a value).  Word k of the group is
// at p[k * ws]: global states are contiguous (ws = 1), LDS states are
// stored word-major (ws = the cache size), so the lanes of a wave that
// update random groups spread over all LDS banks.  In LDS a MIN/MAX word is
// read first and the atomic issued only if the row improves it: the read
// may be stale, which costs at most an unneeded (monotone) atomic.
template <bool LDS, int NS>
__device__ __forceinline__ void update_vals(const GroupArgs& a, uint64_t* p, uint32_t ws,
                                            const uint64_t (&vs)[NS]) {
  atomicAdd((unsigned long long*)&p[0], 1ull);
#pragma unroll
  for (uint32_t i = 0; i < MTL_GROUP_MAX_AGGS; ++i) {
    if (i >= a.n_aggs) break;
    const AggSlot& s = a.aggs[i];
    if (s.dup || s.kind == K_COUNT) continue;
    uint64_t* w = p + (uint64_t)s.word * ws;
    const uint64_t v = pick_src<NS>(vs, s.src);
    switch (s.kind) {
      case K_SUMI: {
        int64_t lo, hi;
        split_i64((int64_t)v, lo, hi);
        atomicAdd((unsigned long long*)w, (unsigned long long)lo);
        if (hi) atomicAdd((unsigned long long*)(w + ws), (unsigned long long)hi);
        break;
      }
      case K_SUMF: unsafeAtomicAdd((double*)w, __longlong_as_double((long long)v)); break;
      case K_MINI:
        if (!LDS || (long long)v < (long long)lds_peek(w)) atomicMin((long long*)w, (long long)v);
        break;
      case K_MAXI:
        if (!LDS || (long long)v > (long long)lds_peek(w)) atomicMax((long long*)w, (long long)v);
        break;
      case K_MINF:
        if (!LDS || v < lds_peek(w)) atomicMin((unsigned long long*)w, (unsigned long long)v);
        break;
      case K_MAXF:
        if (!LDS || v > lds_peek(w)) atomicMax((unsigned long long*)w, (unsigned long long)v);
        break;
      default: break;
    }
  }
}

// Merge a partial state (an LDS cache slot, word k at p[k * ws]) into a
// global group.
__device__ __forceinline__ void merge_state(const GroupArgs& a, uint64_t* g, const uint64_t* p,
                                            uint32_t ws) {
  atomicAdd((unsigned long long*)&g[0], (unsigned long long)p[0]);
  for (uint32_t i = 0; i < a.n_aggs; ++i) {
    const AggSlot& s = a.aggs[i];
    if (s.dup) continue;
    uint64_t* w = g + s.word;
    const uint64_t* v = p + (uint64_t)s.word * ws;
    switch (s.kind) {
      case K_SUMI:
        atomicAdd((unsigned long long*)w, (unsigned long long)v[0]);
        if (v[ws]) atomicAdd((unsigned long long*)(w + 1), (unsigned long long)v[ws]);
        break;
      case K_SUMF: unsafeAtomicAdd((double*)w, __longlong_as_double((long long)v[0])); break;
      case K_MINI: atomicMin((long long*)w, (long long)v[0]); break;
      case K_MAXI: atomicMax((long long*)w, (long long)v[0]); break;
      case K_MINF: atomicMin((unsigned long long*)w, (unsigned long long)v[0]); break;
      case K_MAXF: atomicMax((unsigned long long*)w, (unsigned long long)v[0]); break;
      default: break;
    }
  }
}

__device__ __forceinline__ uint64_t global_slot(const GroupArgs& a, uint64_t key) {
  if (a.dense) return key;
  if (key == GEMPTY) return a.cap;
  uint64_t mask = a.cap - 1, s = gfmix(key) & mask;
  for (uint64_t p = 0; p < a.cap; ++p) {
    uint64_t k = a.tkeys[s];
    if (k == key) return s;
    if (k == GEMPTY) {
      unsigned long long old = atomicCAS((unsigned long long*)&a.tkeys[s], GEMPTY, key);
      if (old == GEMPTY || old == key) return s;
    }
    s = (s + 1) & mask;
  }
  atomicOr(a.err, 64u);
  return GEMPTY;
}

// Table initialisation; block 0 also clears the 64 error words (a.err).
__global__ void k_group_init(GroupArgs a, uint64_t n_groups) {
  if (blockIdx.x == 0 && threadIdx.x < 64) a.err[threadIdx.x] = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_groups;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t* p = a.st + i * a.W;
    for (uint32_t w = 0; w < a.W; ++w) p[w] = a.init[w];
    if (!a.dense) a.tkeys[i] = GEMPTY;
  }
}

// The cache of the spill mode: 32-bit keys in sets of kWays consecutive
// slots (set = multiplicative hash of the key), two 16-byte LDS reads per
// lookup.  A key is inserted into the first empty way of its set (after
// the doorkeeper, when on) with a CAS; ways are only ever filled, so a key
// holds at most one way and a stale read costs at most a failed CAS.  The
// slot, or -1 when the row is not cached (its set is full of other keys,
// or the doorkeeper has not seen it before).
constexpr uint32_t kEmpty32 = 0xFFFFFFFFu;
constexpr uint32_t kWays = 8;
__device__ __forceinline__ int set_lookup(uint32_t* ck32, uint32_t cs, uint32_t* door,
                                          uint32_t use_door, uint32_t k) {
  const uint32_t sets = cs / kWays;  // a power of two, at least 16 (cache of >= 128 slots)
  const uint32_t s0 = (k * 0x9E3779B1u) >> (32 - (31 - __clz((int)sets)));
  uint32_t* set = ck32 + s0 * kWays;
  const uint4 w0 = ((const uint4*)set)[0], w1 = ((const uint4*)set)[1];
  const uint32_t w[kWays] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
  int first_empty = -1;
#pragma unroll
  for (int i = kWays - 1; i >= 0; --i) {
    if (w[i] == k) return (int)(s0 * kWays + i);
    if (w[i] == kEmpty32) first_empty = i;
  }
  if (first_empty < 0) return -1;
  if (use_door) {
    // doorkeeper: a group is cached on its second sighting in this block,
    // so the one-off keys of a long tail do not take the ways of hot groups
    const uint32_t b = (k * 0x85EBCA6Bu) >> (32 - 16);
    const uint32_t was = atomicOr(&door[b >> 5], 1u << (b & 31));
    if (!((was >> (b & 31)) & 1u)) return -1;
  }
  for (uint32_t i = (uint32_t)first_empty; i < kWays; ++i) {
    const uint32_t old = atomicCAS(&set[i], kEmpty32, k);
    if (old == kEmpty32 || old == k) return (int)(s0 * kWays + i);
  }
  return -1;
}

// The LDS cache updates of a round of GE rows, aggregate-major: the kind of
// each update is decided once per round, not once per row.  Word k of cache
// slot i is at cst[k * ws + i]; rows with slot < 0 are skipped, or, with
// SPARE, sent to the spare column ws - 1 (never read back), which avoids a
// branch per row where skipped rows are rare.
template <int NS, int GE_, bool SPARE>
__device__ __forceinline__ void cache_update(const GroupArgs& a, uint64_t* cst, uint32_t ws,
                                             const int (&slot)[GE_],
                                             const uint64_t (&v)[GE_][NS]) {
  int sl[GE_];
  bool on[GE_];
#pragma unroll
  for (int e = 0; e < GE_; ++e) {
    on[e] = SPARE || slot[e] >= 0;
    sl[e] = slot[e] >= 0 ? slot[e] : (int)ws - 1;
  }
#pragma unroll
  for (int e = 0; e < GE_; ++e)
    if (on[e]) atomicAdd((unsigned long long*)&cst[sl[e]], 1ull);
  // sources outer (compile-time: each op reads its source's registers
  // directly), the source's ops inner
#pragma unroll
  for (uint32_t j = 0; j < (uint32_t)NS; ++j) {
    if (j >= a.n_src) break;
#pragma unroll
    for (uint32_t o = 0; o < MTL_GROUP_MAX_AGGS; ++o) {
      if (a.op_beg[j] + o >= a.op_beg[j + 1]) break;
      const uint32_t op = a.ops[a.op_beg[j] + o], kind = op & 15u;
      uint64_t* w = cst + (uint64_t)((op >> 4) & 255u) * ws;
      switch (kind) {
        case K_SUMI:
#pragma unroll
          for (int e = 0; e < GE_; ++e) {
            int64_t lo, hi;
            split_i64((int64_t)v[e][j], lo, hi);
            if (on[e]) {
              atomicAdd((unsigned long long*)&w[sl[e]], (unsigned long long)lo);
              if (hi) atomicAdd((unsigned long long*)&w[ws + sl[e]], (unsigned long long)hi);
            }
          }
          break;
        case K_SUMF:
#pragma unroll
          for (int e = 0; e < GE_; ++e)
            if (on[e]) unsafeAtomicAdd((double*)&w[sl[e]], __longlong_as_double((long long)v[e][j]));
          break;
        case K_MINI:
#pragma unroll
          for (int e = 0; e < GE_; ++e) {
            const long long x = (long long)v[e][j];
            if (on[e] && x < (long long)lds_peek(&w[sl[e]])) atomicMin((long long*)&w[sl[e]], x);
          }
          break;
        case K_MAXI:
#pragma unroll
          for (int e = 0; e < GE_; ++e) {
            const long long x = (long long)v[e][j];
            if (on[e] && x > (long long)lds_peek(&w[sl[e]])) atomicMax((long long*)&w[sl[e]], x);
          }
          break;
        case K_MINF:
#pragma unroll
          for (int e = 0; e < GE_; ++e) {
            const uint64_t x = v[e][j];
            if (on[e] && x < lds_peek(&w[sl[e]]))
              atomicMin((unsigned long long*)&w[sl[e]], (unsigned long long)x);
          }
          break;
        case K_MAXF:
#pragma unroll
          for (int e = 0; e < GE_; ++e) {
            const uint64_t x = v[e][j];
            if (on[e] && x > lds_peek(&w[sl[e]]))
              atomicMax((unsigned long long*)&w[sl[e]], (unsigned long long)x);
          }
          break;
        default: break;
      }
    }
  }
}

// 12 waves per block: two blocks (and their LDS caches) per CU at 6 waves
// per SIMD, within the SGPR-bound limit of 7.
constexpr int kAggBlock = 768;
constexpr uint32_t kDoorBits = 1u << 16;  // doorkeeper bitset of the hashed cache (8 KB)
static_assert(kDoorBits == 1u << 16, "set_lookup takes 16 bits of hash for the doorkeeper");
constexpr int GE = 4;  // rows per thread per round, loaded together

// NS: compile-time bound on the number of value sources (registers for the
// values of GE rows are sized by it, not by the aggregate count).
// Specialised variants fix the shape at compile time: KW != 0 means no
// selection vector and one packed key column of KW bytes, SW0 / SW1 the
// byte widths of the two sources, MODE 1 the directly indexed cache and
// MODE 2 the hashed cache with spilling (0: all of it read from the
// arguments).  With no runtime branch around the column loads, the loads of
// a round go out back to back and stay in flight across the updates.
template <int NS, int KW = 0, int SW0 = 0, int SW1 = 0, int MODE = 0>
__global__ __launch_bounds__(kAggBlock, NS <= 4 ? 6 : 4) void k_group_agg(GroupArgs a) {
  static_assert(KW == 0 || NS == 2, "specialised variants take two sources");
  const uint32_t* sel = KW ? nullptr : a.sel;
  const bool direct = MODE ? MODE == 1 : a.cache_direct != 0;
  const bool spilling = MODE ? MODE == 2 : a.spill != 0;
  extern __shared__ uint64_t smem[];
  __shared__ uint32_t s_spill;
  const uint32_t cs = a.cache_slots;
  const uint32_t ws = cs + 1;  // word stride of the cache states: cs slots and a spare
  uint64_t* ckeys = smem;
  uint64_t* cst = smem + cs;
  uint32_t* door = (uint32_t*)(cst + (uint64_t)ws * a.W);  // kDoorBits bits when a.door
  if (threadIdx.x == 0) s_spill = 0;
  if (a.door)
    for (uint32_t i = threadIdx.x; i < kDoorBits / 32; i += blockDim.x) door[i] = 0;
  // spill mode: packed keys below 2^32 - 1 in a set-associative cache of
  // 32-bit keys (the first half of the key region)
  uint32_t* ck32 = (uint32_t*)ckeys;
  for (uint32_t i = threadIdx.x; i < ws; i += blockDim.x) {
    if (i < cs) {
      if (MODE == 2) ck32[i] = kEmpty32;
      else ckeys[i] = direct ? (uint64_t)i : GEMPTY;
    }
    for (uint32_t w = 0; w < a.W; ++w) cst[(uint64_t)w * ws + i] = a.init[w];
  }
  __syncthreads();
  const uint64_t n = sel ? dev_count(a.sel_count, a.sel_cap) : a.n_rows;
  const uint64_t per = (n + gridDim.x - 1) / gridDim.x;
  const uint64_t start = (uint64_t)blockIdx.x * per;
  const uint64_t end = start + per < n ? start + per : n;
  const uint64_t step = (uint64_t)blockDim.x * GE;
  // software pipeline: the raw bits of the first key column and of every
  // source for the next round are loaded while this round updates the
  // cache, so the column loads of a wave stay in flight behind its atomics
  // (rows past the end read row 0)
  uint64_t nkb[GE], nvb[NS][GE];  // general variants
  typename RawT<KW>::type fk[GE];  // specialised variants
  typename RawT<SW0>::type fv0[GE];
  typename RawT<SW1>::type fv1[GE];
  auto fetch = [&](uint64_t b0) {
    uint64_t r[GE];
#pragma unroll
    for (int e = 0; e < GE; ++e) {
      const uint64_t k = b0 + (uint64_t)e * blockDim.x + threadIdx.x;
      r[e] = k < end ? (sel ? (uint64_t)sel[k] : k) : 0;
    }
    if constexpr (KW != 0) {
      load_raw<GE, KW>(a.key_col[0], r, fk);
      load_raw<GE, SW0>(a.src_col[0], r, fv0);
      load_raw<GE, SW1>(a.src_col[1], r, fv1);
    } else {
      load_bits<GE>(a.key_col[0], a.key_w0, r, nkb);
#pragma unroll
      for (uint32_t j = 0; j < (uint32_t)NS; ++j)
        load_bits<GE>(a.src_col[j], a.src_w[j], r, nvb[j]);
    }
  };
  // with more sources the registers of a second round in flight would cost
  // occupancy: those variants load at the top of each round
  constexpr bool kPipe = NS <= 2;
  if (kPipe && start < end) fetch(start);
  for (uint64_t base = start; base < end; base += step) {
    uint64_t row[GE], key[GE];
    bool ok[GE];
    uint64_t kb[GE], vb[NS][GE];
    if (!kPipe) fetch(base);
#pragma unroll
    for (int e = 0; e < GE; ++e) {
      const uint64_t k = base + (uint64_t)e * blockDim.x + threadIdx.x;
      ok[e] = k < end;
      row[e] = ok[e] ? (sel ? (uint64_t)sel[k] : k) : 0;
      if constexpr (KW != 0) {
        kb[e] = fk[e];
        vb[0][e] = fv0[e];
        vb[NS - 1][e] = fv1[e];
      } else {
        kb[e] = nkb[e];
#pragma unroll
        for (uint32_t j = 0; j < (uint32_t)NS; ++j) vb[j][e] = nvb[j][e];
      }
    }
    if (kPipe && base + step < end) fetch(base + step);
    uint64_t v[GE][NS];
#pragma unroll
    for (int e = 0; e < GE; ++e)
#pragma unroll
      for (uint32_t j = 0; j < (uint32_t)NS; ++j)
        v[e][j] = (ok[e] && j < a.n_src) ? src_bits(a, j, vb[j][e]) : 0ull;
#pragma unroll
    for (int e = 0; e < GE; ++e) {
      key[e] = 0;
      bool in_range;
      if constexpr (KW != 0) {
        // one packed key of KW bytes: its offset from the minimum
        const int64_t kv = KW == 8 ? (int64_t)kb[e] : (int64_t)(int32_t)(uint32_t)kb[e];
        key[e] = key_off(kv, a.key_min[0]);
        in_range = key[e] < a.key_span[0];
      } else {
        in_range = pack_key_bits(a, row[e], kb[e], key[e]);
      }
      if (ok[e] && !in_range) {
        atomicOr(a.err, 8u);
        ok[e] = false;
      }
    }
    int slot[GE];
#pragma unroll
    for (int e = 0; e < GE; ++e) {
      slot[e] = -1;
      if (!ok[e]) {
      } else if (direct) {
        slot[e] = (int)key[e];  // the whole (dense) domain is cached
      } else if constexpr (MODE == 2) {
        slot[e] = set_lookup(ck32, cs, door, a.door, (uint32_t)key[e]);
      } else if (cs && key[e] != GEMPTY) {
        const uint64_t hm = gfmix(key[e]);
        uint32_t h = (uint32_t)hm & (cs - 1);
#pragma unroll 1
        for (int p = 0; p < 8; ++p) {
          uint32_t si = (h + p) & (cs - 1);
          uint64_t ck = __hip_atomic_load(&ckeys[si], __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_WORKGROUP);
          if (ck == key[e]) {
            slot[e] = (int)si;
            break;
          }
          if (ck == GEMPTY) {
            if (a.door) {
              // doorkeeper: a group is cached on its second sighting in this
              // block, so the one-off keys of a long tail do not take the
              // slots of hot groups (first come, first cached otherwise)
              const uint32_t b = (uint32_t)(hm >> 32) & (kDoorBits - 1);
              const uint32_t was = atomicOr(&door[b >> 5], 1u << (b & 31));
              if (!((was >> (b & 31)) & 1u)) break;
            }
            unsigned long long old = atomicCAS((unsigned long long*)&ckeys[si], GEMPTY,
                                               (unsigned long long)key[e]);
            if (old == GEMPTY || old == key[e]) {
              slot[e] = (int)si;
              break;
            }
          }
        }
      }
    }
    cache_update<NS, GE, MODE == 1>(a, cst, ws, slot, v);
#pragma unroll
    for (int e = 0; e < GE; ++e) {
      if constexpr (MODE == 1) break;  // every row in range is cached
      const bool miss = ok[e] && slot[e] < 0;
      if (spilling) {
        // every lane of the wave is here: one LDS atomic per wave reserves
        // the wave's spill entries, in lane order
        const uint64_t mb = __ballot(miss);
        if (mb) {
          const int leader = __ffsll((unsigned long long)mb) - 1;
          uint32_t at = 0;
          if (lane_id() == leader) at = atomicAdd(&s_spill, (uint32_t)__popcll(mb));
          at = __shfl(at, leader, WAVE) + (uint32_t)__popcll(mb & lanemask_lt());
          if (miss) {
            const uint64_t o = start + at;
            a.spill_key[o] = (uint32_t)key[e];
            uint32_t* pay = a.spill_pay + o * a.n_pay;
#pragma unroll
            for (uint32_t j = 0; j < (uint32_t)NS; ++j) {
              if (j >= a.n_src) break;
              const bool wide = KW ? (j == 0 ? SW0 : SW1) == 8 : a.pay_hi[j] != a.pay_lo[j];
              if (wide) {
                pay[a.pay_lo[j]] = (uint32_t)v[e][j];
                pay[a.pay_hi[j]] = (uint32_t)(v[e][j] >> 32);
              } else {
                pay[a.pay_lo[j]] = narrow_src(a, j, v[e][j]);
              }
            }
          }
        }
      } else if (MODE != 1 && miss) {
        uint64_t g = global_slot(a, key[e]);
        if (g != GEMPTY) update_vals<false, NS>(a, a.st + g * a.W, 1, v[e]);
      }
    }
  }
  __syncthreads();
  if (spilling) {
    if (threadIdx.x == 0) a.spill_n[blockIdx.x] = s_spill;
    if (blockIdx.x == 0 && threadIdx.x == 0) a.spill_n[gridDim.x] = (uint32_t)per;
    const uint64_t h0 = (uint64_t)blockIdx.x * cs;
    for (uint32_t i = threadIdx.x; i < cs; i += blockDim.x) {
      const uint64_t key = MODE == 2 ? (ck32[i] == kEmpty32 ? GEMPTY : (uint64_t)ck32[i]) : ckeys[i];
      a.hot_key[h0 + i] = (key == GEMPTY || cst[i] == 0) ? GEMPTY : key;
    }
    // state words slot-major, consecutive threads on consecutive words
    for (uint32_t t = threadIdx.x; t < cs * a.W; t += blockDim.x)
      a.hot_st[h0 * a.W + t] = cst[(uint64_t)(t % a.W) * ws + t / a.W];
    return;
  }
  if (direct && a.dstage) {
    // the cache's state words as they lie in LDS (word-major, stride ws)
    uint64_t* o = a.dstage + (uint64_t)blockIdx.x * ws * a.W;
    for (uint32_t t = threadIdx.x; t < ws * a.W; t += blockDim.x) o[t] = cst[t];
    return;
  }
  for (uint32_t i = threadIdx.x; i < cs; i += blockDim.x) {
    uint64_t key = ckeys[i];
    if (key == GEMPTY || cst[i] == 0) continue;  // word 0: the row count
    uint64_t g = global_slot(a, key);
    if (g != GEMPTY) merge_state(a, a.st + g * a.W, cst + i, ws);
  }
}

// ---------------------------------------------------------------------------
// Spilled rows (cache misses of a large dense domain) are partitioned by
// packed key into ranges of 2^ps keys (partition p = key >> ps), most
// significant digits first, at most 7 bits of the partition number per
// scatter level; then one block per partition aggregates its rows in LDS
// and writes the partition's table entries whole (no table clearing pass,
// no device-scope atomics per row).  The staged LDS caches of k_group_agg
// are merged in last.  Records are structure-of-arrays: the u32 packed key
// and the 64-bit value sources (source j of record i at val[j * stride + i]).
// ---------------------------------------------------------------------------
constexpr int kPartThreads = 1024;  // partition counting, scan and scatter blocks
constexpr int kPartChunk = 4096;    // rows per scatter block
constexpr int kPartLocal = 1024;    // buckets one chunk reserves through LDS
constexpr uint32_t kLocalBits = 10;
constexpr uint32_t kPartHistLds = 16384;  // partitions counted in LDS
constexpr uint64_t kPartAggLds = 64 * 1024;  // LDS state of one partition
constexpr int kPartAggThreads = 512;
constexpr uint32_t kPartAggRows = 65536;  // rows per aggregation block

// Exclusive scan of one value per thread over a block of 64-lane waves;
// s_w holds blockDim.x / 64 words.  Every thread of the block calls it.
__device__ __forceinline__ uint32_t block_scan_excl(uint32_t v, uint32_t* s_w) {
  const uint32_t lane = lane_id(), wave = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, d, WAVE);
    if ((int)lane >= d) x += y;
  }
  if (lane == WAVE - 1) s_w[wave] = x;
  __syncthreads();
  if (wave == 0) {
    uint32_t t = lane < nw ? s_w[lane] : 0u, u = t;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)u, d, WAVE);
      if ((int)lane >= d) u += y;
    }
    if (lane < nw) s_w[lane] = u - t;
  }
  __syncthreads();
  const uint32_t r = x - v + s_w[wave];
  __syncthreads();  // s_w may be reused by the caller
  return r;
}

// The lanes of the wave whose row has the same digit as this lane's
// (digits below 2^bits; lanes without a row match none): one ballot per
// digit bit.  Lets one lane per distinct digit stand for all of them, so a
// skewed digit costs one LDS atomic per wave, not one per lane.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid, uint32_t bits) {
  uint64_t m = __ballot(valid);
  for (uint32_t k = 0; k < bits; ++k) {
    const bool bit = (d >> k) & 1u;
    const uint64_t b = __ballot(valid && bit);
    m &= bit ? b : ~b;
  }
  return valid ? m : 0ull;
}

// This lane's rank among the wave's rows with its digit, after the rows of
// that digit counted in h before (h[d] grows by the wave's count).
__device__ __forceinline__ uint32_t wave_rank(uint32_t* h, uint32_t d, bool valid,
                                              uint32_t bits) {
  const uint64_t m = match_digit(d, valid, bits);
  const int leader = m ? __ffsll((unsigned long long)m) - 1 : 0;
  uint32_t old = 0;
  if (valid && lane_id() == leader) old = atomicAdd(&h[d], (uint32_t)__popcll(m));
  old = (uint32_t)__shfl((int)old, leader, WAVE);
  return old + (uint32_t)__popcll(m & lanemask_lt());
}

// Rows of each partition: one block per spill region (the region of block r
// of k_group_agg starts at r * seg_n[grid] and holds seg_n[r] rows); counted in
// LDS when the partitions fit, else with device atomics per row.
__global__ __launch_bounds__(kPartThreads) void k_part_hist(const uint32_t* key,
                                                            const uint32_t* seg_n, uint32_t ps,
                                                            uint32_t np, uint32_t* count) {
  extern __shared__ uint32_t h[];
  const bool lds = np <= kPartHistLds;
  if (lds)
    for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) h[i] = 0;
  __syncthreads();
  const uint64_t r0 = (uint64_t)blockIdx.x * seg_n[gridDim.x];
  const uint32_t n = seg_n[blockIdx.x];
  for (uint32_t i0 = 0; i0 < n; i0 += 4 * blockDim.x) {
    uint32_t k[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const uint32_t i = i0 + e * blockDim.x + threadIdx.x;
      k[e] = i < n ? key[r0 + i] : 0u;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (i0 + e * blockDim.x + threadIdx.x >= n) continue;
      const uint32_t p = k[e] >> ps;
      if (lds) atomicAdd(&h[p], 1u);
      else atomicAdd(&count[p], 1u);
    }
  }
  __syncthreads();
  if (lds)
    for (uint32_t i = threadIdx.x; i < np; i += blockDim.x)
      if (h[i]) atomicAdd(&count[i], h[i]);
}

// off[p] = rows of the partitions before p, off[np] = all rows; coff the
// same over the aggregation blocks of each partition (one per kPartAggRows
// rows, at least one).  One block.
__device__ __forceinline__ uint32_t part_chunks(uint32_t c) {
  return c ? (c + kPartAggRows - 1) / kPartAggRows : 1u;
}
__global__ __launch_bounds__(kPartThreads) void k_part_scan(const uint32_t* count, uint32_t np,
                                                            uint32_t* off, uint32_t* coff) {
  __shared__ uint32_t s_w[kPartThreads / WAVE];
  const uint32_t per = (np + blockDim.x - 1) / blockDim.x;
  const uint32_t b = threadIdx.x * per, e = b + per < np ? b + per : np;
  uint32_t sum = 0, chunks = 0;
  for (uint32_t i = b; i < e; ++i) {
    sum += count[i];
    chunks += part_chunks(count[i]);
  }
  uint32_t run = block_scan_excl(sum, s_w);
  uint32_t crun = block_scan_excl(chunks, s_w);
  for (uint32_t i = b; i < e; ++i) {
    off[i] = run;
    coff[i] = crun;
    run += count[i];
    crun += part_chunks(count[i]);
  }
  if (threadIdx.x == blockDim.x - 1) {
    off[np] = run;
    coff[np] = crun;
  }
}

// One scatter level: input rows go to the output region of their bucket
// (key >> shift; bucket b's region starts at part_off[b << ushift]) in chunks
// of kPartChunk rows.  A chunk ranks its rows per bucket in LDS (lanes of a
// wave with the same bucket share one atomic), reserves each bucket's run
// with one device atomic, stages the keys and then the payload records in
// LDS in bucket order and writes them with consecutive threads on
// consecutive rows, so each run of a bucket is one stream of records.  The
// first level reads the spill regions (seg_n set); later levels read the
// previous level's output, already grouped by bucket >> coarse, so a chunk
// spans few coarse buckets (buckets past kPartLocal from the chunk's first
// coarse bucket, rare, take one device atomic per row).  Order within a
// bucket is not kept.
struct PartLevel {
  const uint32_t* in_key;
  const uint32_t* in_pay;
  uint32_t* out_key;
  uint32_t* out_pay;
  uint32_t pw;            // payload words per row
  const uint32_t* seg_n;  // segmented input (spill regions: seg_n[n_seg] is
                          // their length), else contiguous
  uint32_t n_seg;
  uint32_t chunks_per_seg;  // bound on the chunks of one region
  uint64_t n_in;
  const uint32_t* part_off;
  uint32_t* used;  // rows placed so far per bucket (cleared before the level)
  uint32_t shift, ushift, coarse, sorted;
};

// A payload record of PW words, moved as one vector access.
template <int PW>
struct alignas(PW == 2 ? 8 : PW == 4 ? 16 : 4) PayRec {
  uint32_t w[PW];
};

// PW: payload words (0 to 4: one vector access per record; kMaxPay: any
// count up to it, word by word, four at a time).
template <int PW>
__global__ __launch_bounds__(kPartThreads) void k_part_scatter(PartLevel a) {
  constexpr bool kAny = PW > 4;
  constexpr int SW = kAny ? 4 : (PW > 0 ? PW : 1);  // payload words staged at a time
  constexpr int R = kPartChunk / kPartThreads;
  static_assert(kPartLocal == kPartThreads, "one local bucket per thread");
  __shared__ uint32_t hist[kPartLocal];
  __shared__ uint32_t res[kPartLocal];
  __shared__ uint32_t dest[kPartChunk];
  __shared__ uint32_t stage[SW][kPartChunk];
  __shared__ uint32_t s_w[kPartThreads / WAVE];
  __shared__ uint32_t s_tot;
  uint64_t c0, cn;
  if (a.seg_n) {
    const uint32_t seg = blockIdx.x / a.chunks_per_seg, c = blockIdx.x % a.chunks_per_seg;
    const uint64_t n = a.seg_n[seg], first = (uint64_t)c * kPartChunk;
    if (first >= n) return;
    c0 = (uint64_t)seg * a.seg_n[a.n_seg] + first;
    cn = n - first < (uint64_t)kPartChunk ? n - first : (uint64_t)kPartChunk;
  } else {
    c0 = (uint64_t)blockIdx.x * kPartChunk;
    if (c0 >= a.n_in) return;
    cn = a.n_in - c0 < (uint64_t)kPartChunk ? a.n_in - c0 : (uint64_t)kPartChunk;
  }
  const uint32_t tid = threadIdx.x;
  const uint32_t base = a.sorted ? ((a.in_key[c0] >> a.shift) >> a.coarse) << a.coarse : 0u;
  hist[tid] = 0;
  uint32_t key[R], loc[R], rank[R];
  PayRec<SW> pay[R];  // fixed payloads: the whole record, loaded with the keys
#pragma unroll
  for (int e = 0; e < R; ++e) {
    const uint32_t i = e * kPartThreads + tid;
    key[e] = i < cn ? a.in_key[c0 + i] : 0u;
    if constexpr (!kAny && PW > 0)
      if (i < cn) pay[e] = ((const PayRec<SW>*)a.in_pay)[c0 + i];
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < R; ++e) {
    const uint32_t i = e * kPartThreads + tid;
    loc[e] = (key[e] >> a.shift) - base;
    const bool local = i < cn && loc[e] < (uint32_t)kPartLocal;
    rank[e] = wave_rank(hist, local ? loc[e] : 0u, local, kLocalBits);
    if (!local) rank[e] = ~0u;
  }
  __syncthreads();
  const uint32_t hc = hist[tid];
  const uint32_t off = block_scan_excl(hc, s_w);
  if (hc) res[tid] = a.part_off[(base + tid) << a.ushift] + atomicAdd(&a.used[base + tid], hc);
  hist[tid] = off;
  if (tid == kPartThreads - 1) s_tot = off + hc;
  __syncthreads();
  uint32_t si[R];
#pragma unroll
  for (int e = 0; e < R; ++e) {
    const uint32_t i = e * kPartThreads + tid;
    si[e] = ~0u;
    if (i >= cn) continue;
    if (rank[e] != ~0u) {
      si[e] = hist[loc[e]] + rank[e];
      dest[si[e]] = res[loc[e]] + rank[e];
    } else {
      const uint32_t bk = key[e] >> a.shift;
      const uint32_t d = a.part_off[bk << a.ushift] + atomicAdd(&a.used[bk], 1u);
      a.out_key[d] = key[e];
      for (uint32_t w = 0; w < a.pw; ++w)
        a.out_pay[(uint64_t)d * a.pw + w] = a.in_pay[(c0 + i) * a.pw + w];
    }
  }
  const uint32_t nst = s_tot;  // written before the barrier above
#pragma unroll
  for (int e = 0; e < R; ++e)
    if (si[e] != ~0u) stage[0][si[e]] = key[e];
  __syncthreads();
  for (uint32_t i = tid; i < nst; i += kPartThreads) a.out_key[dest[i]] = stage[0][i];
  if constexpr (!kAny && PW > 0) {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < R; ++e)
      if (si[e] != ~0u)
#pragma unroll
        for (int w = 0; w < PW; ++w) stage[w][si[e]] = pay[e].w[w];
    __syncthreads();
    for (uint32_t i = tid; i < nst; i += kPartThreads) {
      PayRec<PW> r;
#pragma unroll
      for (int w = 0; w < PW; ++w) r.w[w] = stage[w][i];
      ((PayRec<PW>*)a.out_pay)[dest[i]] = r;
    }
  } else if constexpr (kAny) {
    for (uint32_t w0 = 0; w0 < a.pw; w0 += SW) {
      const uint32_t nw = a.pw - w0 < (uint32_t)SW ? a.pw - w0 : (uint32_t)SW;
      __syncthreads();
#pragma unroll
      for (int e = 0; e < R; ++e) {
        if (si[e] == ~0u) continue;
        const uint64_t row = c0 + e * kPartThreads + tid;
        for (uint32_t w = 0; w < nw; ++w) stage[w][si[e]] = a.in_pay[row * a.pw + w0 + w];
      }
      __syncthreads();
      for (uint32_t i = tid; i < nst; i += kPartThreads)
        for (uint32_t w = 0; w < nw; ++w) a.out_pay[(uint64_t)dest[i] * a.pw + w0 + w] = stage[w][i];
    }
  }
}

template <int NS>
__global__ __launch_bounds__(kPartAggThreads) void k_part_agg(GroupArgs a, const uint32_t* pkey,
                                                              const uint32_t* ppay,
                                                              const uint32_t* part_off,
                                                              const uint32_t* coff, uint32_t np,
                                                              uint32_t ps, uint64_t* xkey,
                                                              uint64_t* xst) {
  extern __shared__ uint64_t cst[];
  __shared__ uint32_t s_p;
  const uint32_t bi = blockIdx.x;
  if (bi >= coff[np]) return;
  if (threadIdx.x == 0) {  // the partition of this block: last p with coff[p] <= bi
    uint32_t lo = 0, hi = np - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) / 2;
      if (coff[mid] <= bi) lo = mid;
      else hi = mid - 1;
    }
    s_p = lo;
  }
  const uint32_t ks = 1u << ps, W = a.W, ws = ks + 1;  // ks groups and a spare column
  for (uint32_t t = threadIdx.x; t < ws * W; t += blockDim.x) cst[t] = a.init[t / ws];
  __syncthreads();
  const uint32_t p = s_p, c = bi - coff[p];
  const uint64_t g0 = (uint64_t)p << ps;
  const uint32_t nk = a.cap - g0 < ks ? (uint32_t)(a.cap - g0) : ks;
  const uint32_t pe = part_off[p + 1];
  const uint32_t b = part_off[p] + c * kPartAggRows;
  const uint32_t e = pe - b < kPartAggRows ? pe : b + kPartAggRows;
  constexpr int R = 4;
  // software pipeline: the next round's keys and payload words are loaded
  // while this round updates the LDS copy (rows past the end read row b)
  uint32_t fk[R], nlo[R][NS], nhi[R][NS];
  auto fetch = [&](uint32_t i0) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint32_t i = i0 + r * blockDim.x + threadIdx.x;
      const uint32_t ii = i < e ? i : b;
      fk[r] = pkey[ii];
      const uint32_t* pr = ppay + (uint64_t)ii * a.n_pay;
#pragma unroll
      for (uint32_t j = 0; j < (uint32_t)NS; ++j) {  // unused sources read word 0
        nlo[r][j] = a.n_pay ? pr[a.pay_lo[j]] : 0u;
        nhi[r][j] = a.n_pay ? pr[a.pay_hi[j]] : 0u;
      }
    }
  };
  if (b < e) fetch(b);
  for (uint32_t i0 = b; i0 < e; i0 += R * blockDim.x) {
    int k[R];
    uint64_t v[R][NS];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      k[r] = i0 + r * blockDim.x + threadIdx.x < e ? (int)(fk[r] - (uint32_t)g0) : -1;
#pragma unroll
      for (uint32_t j = 0; j < (uint32_t)NS; ++j)
        v[r][j] = a.pay_hi[j] != a.pay_lo[j] ? ((uint64_t)nhi[r][j] << 32 | nlo[r][j])
                                              : widen_src(a, j, nlo[r][j]);
    }
    if (i0 + R * blockDim.x < e) fetch(i0 + R * blockDim.x);
    cache_update<NS, R, true>(a, cst, ws, k, v);  // rows past the end: spare column
  }
  __syncthreads();
  if (c == 0) {
    uint64_t* out = a.st + g0 * W;
    for (uint32_t t = threadIdx.x; t < nk * W; t += blockDim.x) out[t] = cst[(t % W) * ws + t / W];
  } else {
    const uint64_t x0 = (uint64_t)(bi - p - 1) * ks;
    for (uint32_t i = threadIdx.x; i < ks; i += blockDim.x)
      xkey[x0 + i] = (i < nk && cst[i] != 0) ? g0 + i : GEMPTY;  // word 0: the row count
    for (uint32_t t = threadIdx.x; t < ks * W; t += blockDim.x)
      xst[x0 * W + t] = cst[(t % W) * ws + t / W];
  }
}

// The staged LDS caches of the spill mode, by cache set: block (s, y)
// takes the ways of set s in stage blocks [32y, 32y + 32), one entry per
// thread, and combines them in an LDS table of their distinct keys (a hot
// key sits in the same set of every block, not always in the same way);
// each key is then merged into the table once per block.
constexpr uint32_t kSetMergeStages = 32;
__global__ __launch_bounds__(256) void k_hot_set_merge(GroupArgs a, uint32_t n_stages,
                                                       uint32_t cs) {
  static_assert(kSetMergeStages * kWays == 256, "one entry per thread");
  extern __shared__ uint64_t sm[];
  uint64_t* tk = sm;        // 256 keys
  uint64_t* ts = sm + 256;  // their states, W words each (slot-major)
  const uint32_t W = a.W, t = threadIdx.x;
  tk[t] = GEMPTY;
  for (uint32_t w = 0; w < W; ++w) ts[(uint64_t)t * W + w] = a.init[w];
  const uint32_t stage = blockIdx.y * kSetMergeStages + t / kWays;
  const uint64_t slot = (uint64_t)stage * cs + blockIdx.x * kWays + t % kWays;
  const uint64_t key = stage < n_stages ? a.hot_key[slot] : GEMPTY;
  __syncthreads();
  if (key != GEMPTY) {
    uint32_t h = (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 56);
    for (;; h = (h + 1) & 255u) {  // 256 slots for at most 256 keys: ends
      const unsigned long long old =
          atomicCAS((unsigned long long*)&tk[h], GEMPTY, (unsigned long long)key);
      if (old == GEMPTY || old == key) break;
    }
    merge_state(a, ts + (uint64_t)h * W, a.hot_st + slot * W, 1);
  }
  __syncthreads();
  if (tk[t] != GEMPTY) merge_state(a, a.st + tk[t] * W, ts + (uint64_t)t * W, 1);
}

// Combine partial state q into the register copy r (word k in r[k], q[k];
// WM: compile-time bound on the state words).
template <int WM>
__device__ __forceinline__ void combine_state(const GroupArgs& a, uint64_t (&r)[WM],
                                              const uint64_t (&q)[WM]) {
  r[0] += q[0];
#pragma unroll
  for (uint32_t i = 0; i < MTL_GROUP_MAX_AGGS; ++i) {
    if (i >= a.n_aggs) break;
    const AggSlot& sl = a.aggs[i];
    if (sl.dup || sl.kind == K_COUNT) continue;
    const uint32_t w = sl.word;
    // registers are indexed by a runtime word: select over the (unrolled)
    // candidates instead of indexing the array dynamically
#pragma unroll
    for (uint32_t k = 1; k < (uint32_t)WM; ++k) {
      if (k == w) {
        switch (sl.kind) {
          case K_SUMI: r[k] += q[k]; break;
          case K_SUMF:
            r[k] = (uint64_t)__double_as_longlong(__longlong_as_double((long long)r[k]) +
                                                  __longlong_as_double((long long)q[k]));
            break;
          case K_MINI: r[k] = (int64_t)q[k] < (int64_t)r[k] ? q[k] : r[k]; break;
          case K_MAXI: r[k] = (int64_t)q[k] > (int64_t)r[k] ? q[k] : r[k]; break;
          case K_MINF: r[k] = q[k] < r[k] ? q[k] : r[k]; break;
          case K_MAXF: r[k] = q[k] > r[k] ? q[k] : r[k]; break;
          default: break;
        }
      }
      if (sl.kind == K_SUMI && k == w + 1) r[k] += q[k];
    }
  }
}

// merge_state from a register copy (word k in r[k]).
template <int WM>
__device__ __forceinline__ void merge_regs(const GroupArgs& a, uint64_t* g,
                                           const uint64_t (&r)[WM]) {
  atomicAdd((unsigned long long*)&g[0], (unsigned long long)r[0]);
#pragma unroll
  for (uint32_t i = 0; i < MTL_GROUP_MAX_AGGS; ++i) {
    if (i >= a.n_aggs) break;
    const AggSlot& sl = a.aggs[i];
    if (sl.dup || sl.kind == K_COUNT) continue;
    uint64_t v0 = 0, v1 = 0;
#pragma unroll
    for (uint32_t k = 1; k < (uint32_t)WM; ++k) {
      if (k == sl.word) v0 = r[k];
      if (k == sl.word + 1) v1 = r[k];
    }
    uint64_t* w = g + sl.word;
    switch (sl.kind) {
      case K_SUMI:
        atomicAdd((unsigned long long*)w, (unsigned long long)v0);
        if (v1) atomicAdd((unsigned long long*)(w + 1), (unsigned long long)v1);
        break;
      case K_SUMF: unsafeAtomicAdd((double*)w, __longlong_as_double((long long)v0)); break;
      case K_MINI: atomicMin((long long*)w, (long long)v0); break;
      case K_MAXI: atomicMax((long long*)w, (long long)v0); break;
      case K_MINF: atomicMin((unsigned long long*)w, (unsigned long long)v0); break;
      case K_MAXF: atomicMax((unsigned long long*)w, (unsigned long long)v0); break;
      default: break;
    }
  }
}

// Staged partial states (slot i of stage block s: key skey[s * ks + i],
// GEMPTY when unused, W words at sst[(s * ks + i) * W]) into the finished
// table.  Block (x, y) takes stage blocks [y * G, y * G + G) and one thread a
// slot column: runs of one key down the column are combined in registers
// and merged with one set of device atomics, so a group staged by many
// blocks (a hot key in every LDS cache, a large partition's extra blocks)
// costs G times fewer contended atomics.  The next stage's key and words
// are loaded before the current one is combined or merged.  With coff set,
// only stage blocks below coff[np] - np are live.  WM: compile-time bound
// on the state words.
template <int WM>
__global__ __launch_bounds__(256) void k_stage_reduce(GroupArgs a, const uint64_t* skey,
                                                      const uint64_t* sst, uint64_t n_stages,
                                                      uint32_t ks, uint32_t G,
                                                      const uint32_t* coff, uint32_t np) {
  const uint64_t live = coff ? (uint64_t)(coff[np] - np) : n_stages;
  const uint64_t s0 = (uint64_t)blockIdx.y * G;
  if (s0 >= live) return;
  const uint64_t s1 = s0 + G < live ? s0 + G : live;
  const uint32_t W = a.W;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ks; i += gridDim.x * blockDim.x) {
    uint64_t cur = GEMPTY;
    uint64_t r[WM], nq[WM];
    uint64_t nk = skey[s0 * ks + i];
#pragma unroll
    for (uint32_t k = 0; k < (uint32_t)WM; ++k) nq[k] = k < W ? sst[(s0 * ks + i) * W + k] : 0ull;
    for (uint64_t s = s0; s < s1; ++s) {
      const uint64_t key = nk;
      uint64_t q[WM];
#pragma unroll
      for (uint32_t k = 0; k < (uint32_t)WM; ++k) q[k] = nq[k];
      if (s + 1 < s1) {
        const uint64_t slot = (s + 1) * ks + i;
        nk = skey[slot];
#pragma unroll
        for (uint32_t k = 0; k < (uint32_t)WM; ++k) nq[k] = k < W ? sst[slot * W + k] : 0ull;
      }
      if (key == GEMPTY) continue;
      if (key != cur) {
        if (cur != GEMPTY) merge_regs<WM>(a, a.st + cur * W, r);
        cur = key;
#pragma unroll
        for (uint32_t k = 0; k < (uint32_t)WM; ++k) r[k] = q[k];
      } else {
        combine_state<WM>(a, r, q);
      }
    }
    if (cur != GEMPTY) merge_regs<WM>(a, a.st + cur * W, r);
  }
}

// The staged caches of a directly indexed domain (k_group_agg with dstage:
// word k of slot i of block b at (b * W + k) * (cs + 1) + i) into the
// table.  Block (x, y) takes stage blocks [y * kDenseReduceG, ...) and one
// thread a (word, slot) pair: its kDenseReduceG partial values are loaded
// at once (consecutive threads, consecutive words), combined by the word's
// operation (dense_word_ops) and merged with one device atomic, so a word
// costs kDenseReduceG times fewer atomics than a merge per block.
constexpr uint32_t kDenseReduceG = 32;
enum : uint8_t { DW_NONE, DW_ADD, DW_ADDF, DW_MINI, DW_MAXI, DW_MINU, DW_MAXU };
struct DenseWordOps {
  uint8_t op[kMaxW];
};

__device__ __forceinline__ uint64_t dense_combine(uint8_t op, uint64_t r, uint64_t q) {
  switch (op) {
    case DW_ADD: return r + q;
    case DW_ADDF:
      return (uint64_t)__double_as_longlong(__longlong_as_double((long long)r) +
                                            __longlong_as_double((long long)q));
    case DW_MINI: return (int64_t)q < (int64_t)r ? q : r;
    case DW_MAXI: return (int64_t)q > (int64_t)r ? q : r;
    case DW_MINU: return q < r ? q : r;
    case DW_MAXU: return q > r ? q : r;
    default: return r;
  }
}

__global__ __launch_bounds__(256) void k_dense_reduce(GroupArgs a, DenseWordOps ops,
                                                      const uint64_t* stg, uint32_t nb,
                                                      uint32_t cs) {
  const uint32_t ws = cs + 1, W = a.W;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= W * ws) return;
  const uint32_t k = t / ws, i = t % ws;
  if (i >= cs) return;  // the spare slot
  const uint8_t op = ops.op[k];
  if (op == DW_NONE) return;
  const uint32_t b0 = blockIdx.y * kDenseReduceG;
  const uint32_t nbk = nb - b0 < kDenseReduceG ? nb - b0 : kDenseReduceG;
  uint64_t q[kDenseReduceG];
#pragma unroll
  for (uint32_t u = 0; u < kDenseReduceG; ++u)
    q[u] = u < nbk ? stg[((uint64_t)(b0 + u) * W) * ws + t] : 0ull;
  const uint64_t id = a.init[k];  // the identity of the word's operation
  uint64_t r = id;
#pragma unroll
  for (uint32_t u = 0; u < kDenseReduceG; ++u)
    if (u < nbk) r = dense_combine(op, r, q[u]);
  if (r == id) return;  // merging the identity changes nothing
  unsigned long long* w = (unsigned long long*)&a.st[global_slot(a, i) * W + k];
  switch (op) {
    case DW_ADD: atomicAdd(w, (unsigned long long)r); break;
    case DW_ADDF: unsafeAtomicAdd((double*)w, __longlong_as_double((long long)r)); break;
    case DW_MINI: atomicMin((long long*)w, (long long)r); break;
    case DW_MAXI: atomicMax((long long*)w, (long long)r); break;
    case DW_MINU: atomicMin(w, (unsigned long long)r); break;
    default: atomicMax(w, (unsigned long long)r); break;
  }
}

// The operation that combines each state word of a (the count, then each
// non-duplicate aggregate's one or two words), as merge_state applies them.
static DenseWordOps dense_word_ops(const GroupArgs& a) {
  DenseWordOps o{};
  o.op[0] = DW_ADD;
  for (uint32_t i = 0; i < a.n_aggs; ++i) {
    const AggSlot& sl = a.aggs[i];
    if (sl.dup || sl.kind == K_COUNT) continue;
    const uint32_t w = sl.word;
    switch (sl.kind) {
      case K_SUMI: o.op[w] = DW_ADD; o.op[w + 1] = DW_ADD; break;
      case K_SUMF: o.op[w] = DW_ADDF; break;
      case K_MINI: o.op[w] = DW_MINI; break;
      case K_MAXI: o.op[w] = DW_MAXI; break;
      case K_MINF: o.op[w] = DW_MINU; break;
      case K_MAXF: o.op[w] = DW_MAXU; break;
      default: break;
    }
  }
  return o;
}

// Non-empty groups (count word != 0) as a bitmask + per-4096-entry counts;
// launch_compact turns them into ascending group indices.  Thread t of a
// block owns entries 16t .. 16t+15 of the tile.
__global__ __launch_bounds__(256) void k_group_flags(const uint64_t* st, uint32_t W, uint64_t n,
                                                     uint64_t* bm, uint32_t* counts) {
  __shared__ uint32_t s_wave[4];
  const int tid = threadIdx.x, wave = tid / WAVE, lane = tid % WAVE;
  const uint64_t e0 = (uint64_t)blockIdx.x * kTileRows + (uint64_t)tid * 16;
  uint32_t m = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j)
    if (e0 + j < n && st[(e0 + j) * W] != 0) m |= 1u << j;
  uint64_t w = (uint64_t)m << (16 * (tid & 3));
  w |= __shfl_xor(w, 1, WAVE);
  w |= __shfl_xor(w, 2, WAVE);
  if ((tid & 3) == 0) bm[e0 / 64] = w;
  uint32_t c = wave_sum_u32((uint32_t)__popc(m));
  if (lane == 0) s_wave[wave] = c;
  __syncthreads();
  if (tid == 0) counts[blockIdx.x] = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
}

struct EmitArgs {
  const uint32_t* idx;    // group indices
  const uint32_t* count;  // device count of idx
  const uint64_t* st;
  const uint64_t* tkeys;
  uint32_t W, dense, raw, n_keys;
  int64_t key_min[MTL_GROUP_MAX_KEYS];
  uint64_t key_span[MTL_GROUP_MAX_KEYS];
  uint64_t key_stride[MTL_GROUP_MAX_KEYS];
  uint64_t cap;
  uint32_t n_aggs;
  AggSlot aggs[MTL_GROUP_MAX_AGGS];
  // distinct (column, conversion) value sources of the aggregates: each is
  // loaded once per row (SUM, MIN and MAX of one column share one load)
  uint32_t n_src;
  const void* src_col[MTL_GROUP_MAX_AGGS];
  int32_t src_type[MTL_GROUP_MAX_AGGS];
  uint32_t src_conv[MTL_GROUP_MAX_AGGS];
  int64_t* out_keys;   // n x n_keys
  uint64_t* out_cnt;   // n
  uint64_t* out_aggs;  // n x (2 words per agg)
  uint32_t* out_n;     // non-NULL: *count is copied here (read back with the rows)
};

__device__ __forceinline__ uint64_t group_key_of(const EmitArgs& a, uint32_t g) {
  if (a.dense) return (uint64_t)g;
  return g == a.cap ? GEMPTY : a.tkeys[g];
}

// Output row i of the emission: group g's key columns, count and
// aggregate words.
__device__ __forceinline__ void emit_row(const EmitArgs& a, uint32_t i, uint32_t g) {
  const uint64_t key = group_key_of(a, g);
  {
    for (uint32_t j = 0; j < a.n_keys; ++j) {
      int64_t v;
      if (a.raw) v = (int64_t)key;
      else v = a.key_min[j] + (int64_t)((key / a.key_stride[j]) % a.key_span[j]);
      a.out_keys[(uint64_t)i * a.n_keys + j] = v;
    }
    const uint64_t* p = a.st + (uint64_t)g * a.W;
    a.out_cnt[i] = p[0];
    for (uint32_t k = 0; k < a.n_aggs; ++k) {
      uint64_t* o = a.out_aggs + ((uint64_t)i * a.n_aggs + k) * 2;
      const AggSlot& s = a.aggs[k];
      if (s.kind == K_SUMI) {
        uint64_t lo;
        int64_t hi;
        sum128(p + s.word, lo, hi);
        o[0] = lo;
        o[1] = (uint64_t)hi;
      } else {
        o[0] = s.kind == K_COUNT ? p[0] : p[s.word];
        o[1] = 0;
      }
    }
  }
}

__global__ void k_group_emit(EmitArgs a) {
  const uint32_t n = *a.count;
  if (a.out_n && blockIdx.x == 0 && threadIdx.x == 0) *a.out_n = n;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    emit_row(a, i, a.idx[i]);
}

// Top-K over all table entries: primary = one aggregate (two key words),
// then the packed key ascending.
struct GTopkArgs {
  const uint64_t* st;
  const uint64_t* tkeys;
  uint32_t W, dense, raw;
  uint64_t n;  // table entries
  uint64_t cap;
  uint32_t kind, word;
  int32_t desc;
  int32_t avg;         // order by sum / count
  uint64_t avg_scale;  // AVG(DECIMAL): the count's factor 10^scale (else 1)
  uint32_t k;
  TopkStage stage;    // full keys of every wave-list entry
  TopkLists lists0;   // per-block lists written by phase 1
  uint32_t* out_idx;
  uint32_t* out_n;
};

// The words a group's top-K key needs: count, the ordering aggregate's
// one or two state words, and the stored key (hash table).
struct GRaw {
  uint64_t cnt, w0, w1, key;
};

__device__ __forceinline__ GRaw gtopk_load(const GTopkArgs& a, uint64_t g) {
  const uint64_t* p = a.st + g * a.W;
  GRaw r;
  r.cnt = p[0];
  r.w0 = p[a.word];
  r.w1 = a.kind == K_SUMI ? p[a.word + 1] : 0ull;
  r.key = a.dense ? g : (g == a.cap ? GEMPTY : a.tkeys[g]);
  return r;
}

// DuckDB's AVG over an integer column (IntegerAverageOperationHugeint): the
// 128-bit sum converted to long double (64-bit significand, round to nearest
// even), divided by the long double divisor (the count, times 10^scale for
// a DECIMAL), the quotient rounded to a 64-bit significand and then to
// double.  Emulated here with integer arithmetic, so the top-K key of a
// group is bit for bit the AVG the host emits with x87 long double (the
// two roundings included); div = count * 10^scale < 2^64 is exact.
__device__ __forceinline__ double avg_ld(uint64_t lo, int64_t hi, uint64_t div) {
  const bool neg = hi < 0;
  uint64_t al = lo, ah = (uint64_t)hi;
  if (neg) {  // |sum| as two words
    al = ~lo + 1;
    ah = ~(uint64_t)hi + (al == 0 ? 1 : 0);
  }
  if ((ah | al) == 0) return 0.0;
  // |sum| rounded to a 64-bit significand: m * 2^e
  uint64_t m;
  int e;
  if (ah == 0) {
    m = al;
    e = 0;
  } else {
    const int sh = 64 - __clzll((long long)ah);  // bits of ah, 1..64
    m = sh == 64 ? ah : ((ah << (64 - sh)) | (al >> sh));
    const uint64_t drop = sh == 64 ? al : (al & ((1ull << sh) - 1));
    const uint64_t half = 1ull << (sh - 1);
    e = sh;
    if (drop > half || (drop == half && (m & 1))) {
      if (++m == 0) {
        m = 1ull << 63;
        ++e;
      }
    }
  }
  // quotient m / div to a 64-bit significand, a round bit and a sticky bit
  uint64_t acc = m / div, r = m % div;
  int x = e;
  auto next_bit = [&]() -> uint64_t {
    // r < div < 2^64: 2r may carry out of 64 bits, and is then >= div
    const bool carry = (r >> 63) != 0;
    const uint64_t r2 = r << 1;
    const uint64_t bit = (carry || r2 >= div) ? 1u : 0u;
    r = bit ? r2 - div : r2;
    return bit;
  };
  while (!(acc >> 63)) {
    acc = (acc << 1) | next_bit();
    --x;
  }
  const uint64_t g = next_bit();
  if (g && (r != 0 || (acc & 1))) {
    if (++acc == 0) {
      acc = 1ull << 63;
      ++x;
    }
  }
  // long double -> double: round the 64-bit significand to 53 bits
  uint64_t d = acc >> 11;
  const uint64_t low = acc & 0x7FF;
  if (low > 0x400 || (low == 0x400 && (d & 1))) ++d;
  const double v = ldexp((double)d, x + 11);
  return neg ? -v : v;
}

// Key words: k0 carries the order on its own wherever it can (the lists and
// their bounds compare k0 first, and a tie on k0 reads the other words from
// global memory), k1 and k2 complete it, k3 is the group key (ascending).
// An exact 128-bit SUM is (int64 saturation of the value, high 64 bits,
// low 64 bits): the saturated word orders every sum that fits int64, and the
// two exact words order the rest.
__device__ __forceinline__ void gtopk_keys(const GTopkArgs& a, const GRaw& r, uint64_t k[4]) {
  uint64_t k0, k1 = 0, k2 = 0;
  switch (a.kind) {
    case K_COUNT: k0 = enc_i64((int64_t)r.cnt); break;
    case K_SUMI: {
      const uint64_t w[2] = {r.w0, r.w1};
      uint64_t lo;
      int64_t hi;
      sum128(w, lo, hi);
      if (a.avg) {
        k0 = enc_f64(avg_ld(lo, hi, r.cnt * a.avg_scale));
      } else {
        const bool fits = hi == ((int64_t)lo >> 63);  // hi is lo's sign extension
        const int64_t sat = fits ? (int64_t)lo : (hi < 0 ? INT64_MIN : INT64_MAX);
        k0 = enc_i64(sat);
        k1 = enc_i64(hi);
        k2 = lo;
      }
      break;
    }
    case K_SUMF: {
      double v = __longlong_as_double((long long)r.w0);
      k0 = enc_f64(a.avg ? v / (double)r.cnt : v);
      break;
    }
    case K_MINI:
    case K_MAXI: k0 = enc_i64((int64_t)r.w0); break;
    default: k0 = r.w0; break;  // MINF/MAXF already encoded
  }
  if (a.desc) {
    k0 = ~k0;
    k1 = ~k1;
    k2 = ~k2;
  }
  k[0] = k0;
  k[1] = k1;
  k[2] = k2;
  k[3] = a.raw ? enc_i64((int64_t)r.key) : r.key;
}

// Phase 1: GE_TK groups per thread per round, their words loaded together
// (one round trip per round instead of a count load and then a key load).
constexpr int GE_TK = 4;
template <int R>
__global__ __launch_bounds__(TopkShape<R>::NT) void k_group_topk(GTopkArgs a) {
  using S = TopkShape<R>;
  __shared__ TopkBlockLds<S::W, S::CAP> sl;
  topk_block_init(sl);
  WaveTopk<R> wl;
  wl.init(a.k);
  const uint64_t chunk = (a.n + gridDim.x - 1) / gridDim.x;
  const uint64_t start = (uint64_t)blockIdx.x * chunk;
  const uint64_t end = start + chunk < a.n ? start + chunk : a.n;
  for (uint64_t base = start; base < end; base += (uint64_t)S::NT * GE_TK) {
    GRaw r[GE_TK];
#pragma unroll
    for (int e = 0; e < GE_TK; ++e) {
      const uint64_t g = base + (uint64_t)e * S::NT + threadIdx.x;
      r[e] = g < end ? gtopk_load(a, g) : GRaw{0, 0, 0, 0};
    }
    uint64_t kk[GE_TK][4];
    uint64_t m1 = ~0ull, m2 = ~0ull;
#pragma unroll
    for (int e = 0; e < GE_TK; ++e) {
      const uint64_t g = base + (uint64_t)e * S::NT + threadIdx.x;
#pragma unroll
      for (int f = 0; f < 4; ++f) kk[e][f] = 0;
      if (g < end && r[e].cnt != 0) {
        gtopk_keys(a, r[e], kk[e]);
        topk_min2(kk[e][0], m1, m2);
      }
    }
    if (base == start) {  // the bound from the first pass: the block's exact K-th key
      __shared__ TopkSelectLds sq;
      uint64_t k0s[GE_TK];
      bool val[GE_TK];
#pragma unroll
      for (int e = 0; e < GE_TK; ++e) {
        k0s[e] = kk[e][0];
        val[e] = base + (uint64_t)e * S::NT + threadIdx.x < end && r[e].cnt != 0;
      }
      topk_seed_select<S::NT>(k0s, val, a.k, sq, sl);
    }
    const uint64_t bthr =
        __hip_atomic_load(&sl.thr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
    for (int e = 0; e < GE_TK; ++e) {
      const uint64_t g = base + (uint64_t)e * S::NT + threadIdx.x;
      const bool pre = g < end && r[e].cnt != 0 && kk[e][0] <= wl.thr0 && kk[e][0] <= bthr;
      wl.offer(pre, kk[e], (uint32_t)g);
    }
    topk_share_thr(wl, sl);
  }
  topk_block_finish<S::NT>(wl, sl, a.stage, a.lists0);
}

// Top-K of a table of at most kSmallTopk entries in one launch, emission
// included (in place of the per-block lists, their merge levels and the
// emission kernel): every block builds the keys of all non-empty entries in
// LDS and ranks its own 64 entries against them, by the key words in order
// and then the entry index, so ranks are distinct; each of the block's four
// waves counts over a quarter of the entries.  The entry of rank r < K is
// emitted as output row r, and block 0 writes min(K, non-empty entries) to
// *em.out_n.
constexpr uint32_t kSmallTopk = 2048;
constexpr int kSmallTopkThreads = 256;
constexpr uint32_t kSmallTopkRank = 64;  // entries ranked per block

__global__ __launch_bounds__(kSmallTopkThreads) void k_group_topk_small(GTopkArgs a,
                                                                         EmitArgs em) {
  constexpr int PER = kSmallTopk / kSmallTopkThreads;  // entries loaded per thread
  constexpr int NW = kSmallTopkThreads / WAVE;
  __shared__ uint64_t lk[4][kSmallTopk];
  __shared__ uint32_t lidx[kSmallTopk];  // entry of each compacted position
  __shared__ uint32_t lpos[kSmallTopk];  // compacted position of each entry
  __shared__ uint32_t lpart[2][NW][WAVE];
  __shared__ uint32_t ln;
  if (threadIdx.x == 0) ln = 0;
  const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  const uint32_t n_ent = (uint32_t)a.n;
  // every entry's words in flight at once
  GRaw r[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const uint32_t e = threadIdx.x + (uint32_t)i * kSmallTopkThreads;
    r[i] = e < n_ent ? gtopk_load(a, e) : GRaw{0, 0, 0, 0};
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const uint32_t e = threadIdx.x + (uint32_t)i * kSmallTopkThreads;
    const bool v = e < n_ent && r[i].cnt != 0;
    uint64_t kk[4] = {0, 0, 0, 0};
    if (v) gtopk_keys(a, r[i], kk);
    // one LDS reservation per wave: the lanes' positions from the ballot
    const uint64_t bal = __ballot(v);
    uint32_t base = 0;
    if (lane == 0 && bal) base = atomicAdd(&ln, (uint32_t)__popcll(bal));
    base = (uint32_t)__shfl((int)base, 0, WAVE);
    if (v) {
      const uint32_t p = base + (uint32_t)__popcll(bal & ((1ull << lane) - 1));
#pragma unroll
      for (int f = 0; f < 4; ++f) lk[f][p] = kk[f];
      lidx[p] = e;
      lpos[e] = p;
    } else if (e < n_ent) {
      lpos[e] = 0xFFFFFFFFu;
    }
  }
  __syncthreads();
  const uint32_t n = ln;
  if (blockIdx.x == 0 && threadIdx.x == 0) *em.out_n = n < a.k ? n : a.k;
  const uint32_t mine_e = blockIdx.x * kSmallTopkRank + (uint32_t)lane;
  const uint32_t mp = mine_e < n_ent ? lpos[mine_e] : 0xFFFFFFFFu;
  uint32_t below = 0, tie = 0;
  if (mp != 0xFFFFFFFFu) {
    const uint64_t m0 = lk[0][mp];
    // this wave's quarter of the compacted entries, sixteen LDS reads in
    // flight per step
    const uint32_t q = (n + NW - 1) / NW;
    const uint32_t j0 = (uint32_t)wave * q, j1 = j0 + q < n ? j0 + q : n;
    uint32_t j = j0;
    for (; j + 16 <= j1; j += 16) {
      uint64_t x[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) x[u] = lk[0][j + u];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        below += x[u] < m0 ? 1u : 0u;
        tie += x[u] == m0 ? 1u : 0u;
      }
    }
    for (; j < j1; ++j) {
      const uint64_t x = lk[0][j];
      below += x < m0 ? 1u : 0u;
      tie += x == m0 ? 1u : 0u;
    }
  }
  lpart[0][wave][lane] = below;
  lpart[1][wave][lane] = tie;
  __syncthreads();
  if (wave != 0 || mp == 0xFFFFFFFFu) return;
  below = 0;
  tie = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    below += lpart[0][w][lane];
    tie += lpart[1][w][lane];
  }
  if (tie > 1) {  // equal first words (self included): the other words decide
    uint64_t mk[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) mk[f] = lk[f][mp];
    for (uint32_t j = 0; j < n; ++j) {
      if (lk[0][j] != mk[0]) continue;
      bool lt = false;
      bool eq = true;
#pragma unroll
      for (int f = 1; f < 4; ++f) {
        const uint64_t x = lk[f][j];
        if (eq && x != mk[f]) {
          lt = x < mk[f];
          eq = false;
        }
      }
      if (eq) lt = lidx[j] < mine_e;
      below += lt ? 1u : 0u;
    }
  }
  if (below < a.k) emit_row(em, below, mine_e);
}

// Merge level: block b merges lists [b*F, b*F+F) of `in`; the single block
// of the final level writes the table indices of the top K.
__global__ __launch_bounds__(TOPK_MERGE_THREADS) void k_group_topk_merge(GTopkArgs a,
                                                                         TopkLists in,
                                                                         TopkLists out,
                                                                         uint32_t L,
                                                                         uint32_t final_level) {
  __shared__ TopkMergeLds s;
  const int cur = topk_merge_lists(s, in, L, a.k, a.stage);
  if (!final_level) {
    topk_write_list(s, cur, out, a.k);
    return;
  }
  const uint32_t nn = s.n[cur][0];
  for (uint32_t i = threadIdx.x; i < nn; i += TOPK_MERGE_THREADS)
    a.out_idx[i] = a.stage.cid[s.g[cur][i]];
  if (threadIdx.x == 0) *a.out_n = nn;
}

}  // namespace mtl

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct mtl_groups {
  mtl_stream* s = nullptr;
  mtl::GroupArgs args{};
  uint64_t entries = 0;  // table entries (dense domain or cap + 1)
  void* keys_mem = nullptr;
  void* st_mem = nullptr;
  uint64_t mem_bytes[2] = {0, 0};  // allocated sizes of st_mem, keys_mem
  uint64_t bytes = 0;  // charged to the processing budget
  int32_t agg_func[MTL_GROUP_MAX_AGGS];
  int32_t agg_type[MTL_GROUP_MAX_AGGS];
  int32_t agg_scale[MTL_GROUP_MAX_AGGS];  // DECIMAL digits (AVG divides by 10^scale)
};

// Device memory of `bytes` for slot i (0 state words, 1 hashed keys) of a
// group table on stream s: the stream's parked table memory when it is
// large enough and not over four times too large, else a stream-ordered
// allocation (a parked block that does not fit is returned to the pool).
// Called with s->mu held.
static hipError_t table_mem(mtl_stream* s, int i, uint64_t bytes, void** out, uint64_t* got) {
  void*& p = s->parked_table[i];
  if (p && s->parked_bytes[i] >= bytes && s->parked_bytes[i] / 4 <= bytes) {
    *out = p;
    *got = s->parked_bytes[i];
    p = nullptr;
    s->parked_bytes[i] = 0;
    return hipSuccess;
  }
  if (p) {
    (void)hipFreeAsync(p, s->stream);
    p = nullptr;
    s->parked_bytes[i] = 0;
  }
  *got = bytes;
  return hipMallocAsync(out, bytes, s->stream);
}

namespace mtl {

static bool is_float(int32_t t) { return t == MTL_FLOAT || t == MTL_DOUBLE; }

// 10^s as DuckDB's AverageDecimalBindData holds it: a double, widened.
static long double pow10_ld(int32_t s) {
  double p = 1.0;
  for (int32_t i = 0; i < s; ++i) p *= 10.0;
  return (long double)p;
}

static int32_t group_plan(const mtl_group_desc* d, GroupArgs* a, uint64_t* entries) {
  if (!d) return set_error(MTL_ERR_INVALID_ARGUMENT, "desc is NULL");
  if (d->n_keys < 1 || d->n_keys > MTL_GROUP_MAX_KEYS)
    return set_error(MTL_UNSUPPORTED, "GROUP BY with %u keys", d->n_keys);
  if (d->n_aggs > MTL_GROUP_MAX_AGGS)
    return set_error(MTL_UNSUPPORTED, "more than %d aggregates", MTL_GROUP_MAX_AGGS);
  uint64_t rows = d->sel ? d->sel_capacity : d->n_rows;
  if (rows > 0xFFFFFFFFull) return set_error(MTL_UNSUPPORTED, "more than 2^32-1 rows");
  *a = GroupArgs{};
  a->n_rows = d->n_rows;
  a->sel = d->sel;
  a->sel_count = d->sel ? d->sel_count : nullptr;
  a->sel_cap = d->sel_capacity;
  a->n_keys = d->n_keys;
  // packed domain
  long double dom = 1;
  for (uint32_t i = 0; i < d->n_keys; ++i) {
    if (!d->key_col[i] || !type_is_int(d->key_type[i]))
      return set_error(MTL_UNSUPPORTED, "group key %u must be an integer column", i);
    if (d->key_max[i] < d->key_min[i])
      return set_error(MTL_ERR_INVALID_ARGUMENT, "group key %u: max < min", i);
    a->key_col[i] = d->key_col[i];
    a->key_type[i] = d->key_type[i];
    a->key_min[i] = d->key_min[i];
    a->key_span[i] = key_off(d->key_max[i], d->key_min[i]) + 1;  // 0 = 2^64
    dom *= (a->key_span[i] == 0 ? 18446744073709551616.0L : (long double)a->key_span[i]);
  }
  if (dom > 4611686018427387904.0L) {  // 2^62: cannot pack
    if (d->n_keys != 1)
      return set_error(MTL_UNSUPPORTED, "composite group key domain exceeds 2^62");
    a->raw = 1;
  } else {
    uint64_t stride = 1;
    for (int i = (int)d->n_keys - 1; i >= 0; --i) {
      a->key_stride[i] = stride;
      stride *= a->key_span[i];
    }
  }
  uint64_t D = a->raw ? 0 : (uint64_t)dom;
  // aggregate state words
  uint32_t w = 1;
  a->init[0] = 0;
  for (uint32_t i = 0; i < d->n_aggs; ++i) {
    const mtl_agg_spec& sp = d->aggs[i];
    AggSlot& s = a->aggs[i];
    s.col = sp.values;
    s.type = sp.type;
    bool need_col = sp.func != MTL_AGG_COUNT;
    if (need_col && (!sp.values || !type_width(sp.type) || sp.type == MTL_HUGEINT))
      return set_error(MTL_UNSUPPORTED, "aggregate %u: unsupported value column", i);
    switch (sp.func) {
      case MTL_AGG_COUNT: s.kind = K_COUNT; s.word = 0; break;
      case MTL_AGG_SUM:
      case MTL_AGG_AVG:
        if (is_float(sp.type)) {
          s.kind = K_SUMF; s.word = w; a->init[w++] = 0;
        } else {
          s.kind = K_SUMI; s.word = w; a->init[w++] = 0; a->init[w++] = 0;
        }
        break;
      case MTL_AGG_MIN:
        s.word = w;
        if (is_float(sp.type)) { s.kind = K_MINF; a->init[w++] = ~0ull; }
        else { s.kind = K_MINI; a->init[w++] = (uint64_t)INT64_MAX; }
        break;
      case MTL_AGG_MAX:
        s.word = w;
        if (is_float(sp.type)) { s.kind = K_MAXF; a->init[w++] = 0; }
        else { s.kind = K_MAXI; a->init[w++] = (uint64_t)INT64_MIN; }
        break;
      default:
        return set_error(MTL_UNSUPPORTED, "aggregate function %d", sp.func);
    }
    // SUM(x) and AVG(x), or a repeated aggregate, share one state
    s.dup = 0;
    for (uint32_t j = 0; j < i; ++j) {
      const AggSlot& o = a->aggs[j];
      if (!o.dup && o.kind == s.kind && s.kind != K_COUNT && o.col == s.col && o.type == s.type) {
        w = s.word;  // give back the words just taken
        s.word = o.word;
        s.dup = 1;
        break;
      }
    }
  }
  a->n_aggs = d->n_aggs;
  a->W = w;
  // value sources: one per distinct (column, type, conversion)
  a->n_src = 0;
  for (uint32_t i = 0; i < d->n_aggs; ++i) {
    AggSlot& sl = a->aggs[i];
    sl.src = 0;
    if (sl.kind == K_COUNT) continue;
    const uint32_t conv = sl.kind == K_SUMF                            ? SC_DBL
                          : (sl.kind == K_MINF || sl.kind == K_MAXF) ? SC_ENCF
                                                                     : SC_INT;
    uint32_t j = 0;
    while (j < a->n_src &&
           !(a->src_col[j] == sl.col && a->src_type[j] == sl.type && a->src_conv[j] == conv))
      ++j;
    if (j == a->n_src) {
      a->src_col[j] = sl.col;
      a->src_type[j] = sl.type;
      a->src_conv[j] = conv;
      ++a->n_src;
    }
    sl.src = j;
  }
  // payload words of a spilled row: one or two per source (see GroupArgs)
  a->n_pay = 0;
  for (uint32_t j = 0; j < MTL_GROUP_MAX_AGGS; ++j) {
    a->pay_lo[j] = a->pay_hi[j] = 0;
    if (j >= a->n_src) continue;
    a->pay_lo[j] = a->pay_hi[j] = a->n_pay++;
    if (type_width(a->src_type[j]) == 8) a->pay_hi[j] = a->n_pay++;
  }
  a->n_ops = 0;
  for (uint32_t j = 0; j <= MTL_GROUP_MAX_AGGS; ++j) {
    a->op_beg[j] = a->n_ops;
    if (j == MTL_GROUP_MAX_AGGS) break;
    for (uint32_t i = 0; i < d->n_aggs; ++i) {
      const AggSlot& sl = a->aggs[i];
      if (sl.dup || sl.kind == K_COUNT || sl.src != j) continue;
      a->ops[a->n_ops++] = sl.kind | sl.word << 4 | sl.src << 12;
    }
  }
  a->key_w0 = (uint32_t)type_width(a->key_type[0]);
  for (uint32_t j = 0; j < MTL_GROUP_MAX_AGGS; ++j) {
    if (j >= a->n_src) {  // unused: a valid column, never converted
      a->src_col[j] = a->key_col[0];
      a->src_type[j] = a->key_type[0];
    }
    a->src_w[j] = (uint32_t)type_width(a->src_type[j]);
  }
  if (!a->raw && D <= kDenseMax) {
    a->dense = 1;
    a->cap = D;
    *entries = D;
  } else {
    a->dense = 0;
    // the tuned fill (T_GROUP_LOAD percent of the slots at most)
    uint64_t want = std::min<uint64_t>(rows, a->raw ? rows : D);
    const uint64_t load = (uint64_t)tuned(T_GROUP_LOAD);
    a->cap = 16;
    while (a->cap * load < want * 100) a->cap <<= 1;
    *entries = a->cap + 1;
  }
  // LDS cache of hot groups: <= 60 KB (two 768-thread blocks per CU)
  const uint64_t kCacheBytes = 60 * 1024;
  uint32_t slot_bytes = 8 * (1 + a->W);
  uint32_t cs = 1;
  // the largest power of two of slots that fits (1024 for six state words);
  // first come, first cached, up to 8 probes per key
  while ((uint64_t)cs * 2 * slot_bytes <= kCacheBytes) cs *= 2;
  a->cache_direct = 0;
  if (a->dense && D * slot_bytes <= kCacheBytes) {
    // the whole domain fits: index the cache by key, no probing
    cs = (uint32_t)D;
    a->cache_direct = 1;
  }
  a->cache_slots = cs;
  return MTL_OK;
}

// Spill mode for a directly indexed domain too large for the LDS cache,
// from this many rows (smaller inputs update the global table directly).
constexpr uint64_t kSpillMinRows = 1ull << 20;

// The specialised k_group_agg for the shape of `a`, if it has one (see
// k_group_agg); false: the caller launches the general variant.
template <int KW, int S0, int S1>
static void launch_agg_fixed(const GroupArgs& a, dim3 gd, size_t lds, hipStream_t st) {
  if (a.cache_direct) hipLaunchKernelGGL((k_group_agg<2, KW, S0, S1, 1>), gd, dim3(kAggBlock), lds, st, a);
  else hipLaunchKernelGGL((k_group_agg<2, KW, S0, S1, 2>), gd, dim3(kAggBlock), lds, st, a);
}
static bool launch_agg_fast(const GroupArgs& a, dim3 gd, size_t lds, hipStream_t st) {
  if (a.sel || a.n_keys != 1 || a.raw || a.n_src > 2 || !(a.cache_direct || a.spill)) return false;
  auto fits = [](uint32_t w) { return w == 4 || w == 8; };
  if (!fits(a.key_w0) || !fits(a.src_w[0]) || !fits(a.src_w[1])) return false;
  switch ((a.key_w0 == 8) * 4 + (a.src_w[0] == 8) * 2 + (a.src_w[1] == 8)) {
    case 0: launch_agg_fixed<4, 4, 4>(a, gd, lds, st); break;
    case 1: launch_agg_fixed<4, 4, 8>(a, gd, lds, st); break;
    case 2: launch_agg_fixed<4, 8, 4>(a, gd, lds, st); break;
    case 3: launch_agg_fixed<4, 8, 8>(a, gd, lds, st); break;
    case 4: launch_agg_fixed<8, 4, 4>(a, gd, lds, st); break;
    case 5: launch_agg_fixed<8, 4, 8>(a, gd, lds, st); break;
    case 6: launch_agg_fixed<8, 8, 4>(a, gd, lds, st); break;
    default: launch_agg_fixed<8, 8, 8>(a, gd, lds, st); break;
  }
  return true;
}

// Partitioning and aggregation of the spilled rows, then the merge of the
// staged LDS caches (after k_group_agg on the same stream, which left
// `total` rows in its per-block regions of at most seg_max rows).  Writes every
// table entry: no k_group_init before it.
template <int NS>
static int32_t spill_finish(GroupArgs& a, uint32_t agg_grid, uint64_t seg_max, uint64_t total,
                            mtl_stream* s) {
  if (total > 0xFFFFFFFFull) return set_error(MTL_UNSUPPORTED, "more than 2^32-1 spilled rows");
  Arena& ar = s->arena;
  hipStream_t st = s->stream;
  // partition size: the largest power of two of groups whose state fits
  uint32_t ps = 0;
  while ((2ull << ps) * a.W * 8 <= kPartAggLds) ++ps;
  const uint64_t np64 = (a.cap + (1ull << ps) - 1) >> ps;
  if (np64 > 0xFFFFFFFFull) return set_error(MTL_UNSUPPORTED, "too many group partitions");
  const uint32_t np = (uint32_t)np64;
  uint32_t pb = 0;
  while ((1ull << pb) < np) ++pb;
  // scatter levels of at most the tuned fan-out bits each (T_PART_BITS)
  const uint32_t lbits = (uint32_t)tuned(T_PART_BITS);
  const uint32_t levels = pb == 0 ? 1 : (pb + lbits - 1) / lbits;
  const uint32_t lb = (pb + levels - 1) / levels;
  uint32_t ush[8];  // bucket of level l = partition >> ush[l]
  uint64_t nb[8];
  for (uint32_t l = 0; l < levels; ++l) {
    ush[l] = pb > (l + 1) * lb ? pb - (l + 1) * lb : 0;
    nb[l] = (np + (1ull << ush[l]) - 1) >> ush[l];
  }
  uint64_t used_words = 0;
  for (uint32_t l = 0; l < levels; ++l) used_words += nb[l];
  // aggregation blocks beyond the first of each partition (an upper bound)
  // and their staged states
  const uint64_t ks = 1ull << ps, extra = ceil_div(total, kPartAggRows);
  void *cnt, *off, *coff, *used, *xk = nullptr, *xs = nullptr, *bk = nullptr, *bp = nullptr;
  MTL_TRY(ar.alloc((uint64_t)np * 4, &cnt));
  MTL_TRY(ar.alloc((uint64_t)np * 4 + 4, &off));
  MTL_TRY(ar.alloc((uint64_t)np * 4 + 4, &coff));
  MTL_TRY(ar.alloc(used_words * 4, &used));
  if (extra) {
    MTL_TRY(ar.alloc(extra * ks * 8, &xk));
    MTL_TRY(ar.alloc(extra * ks * a.W * 8, &xs));
  }
  // the second key plane and payload (levels alternate between them and the
  // spill arrays, contiguous from the second level on)
  if (total) {
    MTL_TRY(ar.alloc(total * 4, &bk));
    MTL_TRY(ar.alloc(total * 4 * std::max<uint32_t>(a.n_pay, 1), &bp));
  }
  ZeroList zl;
  zl.add((uint32_t*)cnt, np);
  zl.add((uint32_t*)used, used_words);
  MTL_TRY(launch_zero(zl, st));
  if (total) {
    ProfScope p1(s, "k_part_hist");
    hipLaunchKernelGGL(k_part_hist, dim3(agg_grid), dim3(kPartThreads),
                       np <= kPartHistLds ? (size_t)np * 4 : 0, st, a.spill_key, a.spill_n,
                       ps, np, (uint32_t*)cnt);
    MTL_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_part_scan, dim3(1), dim3(kPartThreads), 0, st, (const uint32_t*)cnt, np,
                     (uint32_t*)off, (uint32_t*)coff);
  MTL_HIP(hipGetLastError());
  const uint32_t* fk = a.spill_key;
  const uint32_t* fp = a.spill_pay;
  uint32_t* used_l = (uint32_t*)used;
  for (uint32_t l = 0; total && l < levels; ++l) {
    PartLevel lv{};
    const bool to_b = (l % 2) == 0;
    lv.in_key = fk;
    lv.in_pay = fp;
    lv.out_key = to_b ? (uint32_t*)bk : a.spill_key;
    lv.out_pay = to_b ? (uint32_t*)bp : a.spill_pay;
    lv.pw = a.n_pay;
    lv.n_in = total;
    lv.part_off = (const uint32_t*)off;
    lv.used = used_l;
    lv.shift = ps + ush[l];
    lv.ushift = ush[l];
    lv.coarse = l ? ush[l - 1] - ush[l] : 0;
    lv.sorted = l ? 1u : 0u;
    uint64_t grid;
    if (l == 0) {
      lv.seg_n = a.spill_n;
      lv.n_seg = agg_grid;
      lv.chunks_per_seg = (uint32_t)ceil_div(seg_max, kPartChunk);
      grid = (uint64_t)agg_grid * lv.chunks_per_seg;
    } else {
      grid = ceil_div(total, kPartChunk);
    }
    {
      ProfScope p2(s, "k_part_scatter");
      const dim3 gd((uint32_t)grid), bd(kPartThreads);
      switch (a.n_pay) {
        case 0: hipLaunchKernelGGL(k_part_scatter<0>, gd, bd, 0, st, lv); break;
        case 1: hipLaunchKernelGGL(k_part_scatter<1>, gd, bd, 0, st, lv); break;
        case 2: hipLaunchKernelGGL(k_part_scatter<2>, gd, bd, 0, st, lv); break;
        case 3: hipLaunchKernelGGL(k_part_scatter<3>, gd, bd, 0, st, lv); break;
        case 4: hipLaunchKernelGGL(k_part_scatter<4>, gd, bd, 0, st, lv); break;
        default: hipLaunchKernelGGL(k_part_scatter<kMaxPay>, gd, bd, 0, st, lv); break;
      }
      MTL_HIP(hipGetLastError());
    }
    fk = lv.out_key;
    fp = lv.out_pay;
    used_l += nb[l];
  }
  {
    ProfScope p3(s, "k_part_agg");
    hipLaunchKernelGGL(k_part_agg<NS>, dim3((uint32_t)(np + extra)), dim3(kPartAggThreads),
                       (size_t)(ks + 1) * a.W * 8, st, a, fk, fp, (const uint32_t*)off,
                       (const uint32_t*)coff, np, ps, (uint64_t*)xk, (uint64_t*)xs);
    MTL_HIP(hipGetLastError());
  }
  {
    ProfScope p4(s, "k_stage_reduce");
    const uint32_t cs = a.cache_slots;
    // extra blocks of large partitions: groups of 16 stage blocks
    const uint32_t gx = 16;
    auto reduce = [&](dim3 gd, const uint64_t* k, const uint64_t* v, uint64_t n, uint32_t w,
                      uint32_t g, const uint32_t* co) {
      if (a.W <= 8) hipLaunchKernelGGL(k_stage_reduce<8>, gd, dim3(256), 0, st, a, k, v, n, w, g, co, np);
      else hipLaunchKernelGGL(k_stage_reduce<kMaxW>, gd, dim3(256), 0, st, a, k, v, n, w, g, co, np);
    };
    hipLaunchKernelGGL(k_hot_set_merge,
                       dim3(cs / kWays, (uint32_t)ceil_div(agg_grid, kSetMergeStages)), dim3(256),
                       (size_t)256 * 8 * (1 + a.W), st, a, agg_grid, cs);
    MTL_HIP(hipGetLastError());
    if (extra) {
      reduce(dim3((uint32_t)ceil_div(ks, 256), (uint32_t)ceil_div(extra, gx)), (const uint64_t*)xk,
             (const uint64_t*)xs, extra, (uint32_t)ks, gx, (const uint32_t*)coff);
      MTL_HIP(hipGetLastError());
    }
  }
  return MTL_OK;
}

// Device bytes a grouped aggregate may take at most: the table, the spill
// arrays of a large dense domain, and the top-K and emission scratch (an
// upper bound for admission; 0 entries when the plan is refused).
// Whether the plan for d indexes its table directly, so that
// mtl_groups_fetch returns the groups ascending by their keys (the first
// key most significant).
bool group_is_dense(const mtl_group_desc& d) {
  GroupArgs a;
  uint64_t entries = 0;
  return group_plan(&d, &a, &entries) == MTL_OK && a.dense;
}

uint64_t group_scratch_bytes(const mtl_group_desc& d) {
  GroupArgs a;
  uint64_t entries = 0;
  if (group_plan(&d, &a, &entries) != MTL_OK) return 1u << 20;
  const uint64_t rows = d.sel ? d.sel_capacity : d.n_rows;
  // staged states: the LDS caches of k_group_agg and the partial states of
  // the extra aggregation blocks of large partitions (<= 64 KB per block)
  const uint64_t hot = (uint64_t)cu_count() * 2 * a.cache_slots * (1 + a.W) * 8 +
                       ceil_div(rows, kPartAggRows) * (kPartAggLds + kPartAggLds / a.W);
  return entries * (a.W + (a.dense ? 0 : 1)) * 8 + rows * 4 * (1 + (uint64_t)a.n_pay) * 2 + hot +
         entries * (16 + 8 * (uint64_t)a.n_aggs) + (64u << 20);
}

}  // namespace mtl

using namespace mtl;

extern "C" {

int32_t mtl_group_agg(const mtl_group_desc* desc, mtl_groups** out, mtl_stream* s) {
  if (!out) return set_error(MTL_ERR_INVALID_ARGUMENT, "out is NULL");
  GroupArgs a;
  uint64_t entries;
  MTL_TRY(group_plan(desc, &a, &entries));
  MTL_TRY(resolve_stream(s, &s));
  std::lock_guard<std::mutex> lk(s->mu);
  mtl_groups* g = new mtl_groups();
  g->s = s;
  g->entries = entries;
  for (uint32_t i = 0; i < desc->n_aggs; ++i) {
    g->agg_func[i] = desc->aggs[i].func;
    g->agg_type[i] = desc->aggs[i].type;
    g->agg_scale[i] = desc->aggs[i].type == MTL_DECIMAL ? desc->aggs[i].scale : 0;
  }
  g->bytes = entries * a.W * 8 + 8 + (a.dense ? 0 : entries * 8);
  if (int32_t cs = charge_processing(g->bytes, "group table")) {
    delete g;
    return cs;
  }
  hipError_t e = table_mem(s, 0, entries * a.W * 8 + 8, &g->st_mem, &g->mem_bytes[0]);
  if (e == hipSuccess && !a.dense) e = table_mem(s, 1, entries * 8, &g->keys_mem, &g->mem_bytes[1]);
  if (e != hipSuccess) {
    if (g->st_mem) (void)hipFreeAsync(g->st_mem, s->stream);
    release_processing(g->bytes);
    delete g;
    return set_error(MTL_ERR_PROCESSING, "processing error: group table of %llu entries: %s",
                     (unsigned long long)entries, hipGetErrorString(e));
  }
  // every error return from here on releases the table
  struct Owner {
    mtl_groups* g;
    ~Owner() {
      if (!g) return;
      (void)hipFreeAsync(g->st_mem, g->s->stream);
      if (g->keys_mem) (void)hipFreeAsync(g->keys_mem, g->s->stream);
      release_processing(g->bytes);
      delete g;
    }
  } owner{g};
  s->arena.reset();
  void* err;
  MTL_TRY(s->arena.alloc(256, &err));
  a.st = (uint64_t*)g->st_mem;
  a.tkeys = (uint64_t*)g->keys_mem;
  a.err = (uint32_t*)err;
  uint64_t rows = desc->sel ? desc->sel_capacity : desc->n_rows;
  // spill mode (a directly indexed domain too large for the LDS cache):
  // the partitioned aggregation of the spilled rows writes every entry
  const bool spill = rows && a.dense && !a.cache_direct && rows >= kSpillMinRows &&
                     a.cache_slots >= 16 * kWays;
  if (spill) {
    ZeroList zl;
    zl.add((uint32_t*)err, 64);
    MTL_TRY(launch_zero(zl, s->stream));
  } else {  // k_group_init also clears the error words
    ProfScope ps(s, "k_group_init");
    uint64_t blocks = std::min<uint64_t>(ceil_div(entries, 256), 8192);
    hipLaunchKernelGGL(k_group_init, dim3((uint32_t)std::max<uint64_t>(blocks, 1)), dim3(256),
                       0, s->stream, a, entries);
  }
  MTL_HIP(hipGetLastError());
  if (rows) {
    a.door = (!a.cache_direct && a.cache_slots && rows >= kSpillMinRows) ? 1u : 0u;
    size_t lds = (size_t)a.cache_slots * 8 + (size_t)(a.cache_slots + 1) * 8 * a.W +
                 (a.door ? kDoorBits / 8 : 0);
    // as many blocks as are resident at once (LDS caches of <= 60 KB in the
    // 160 KB of a CU; at most 2 blocks of 768 threads), so every block gets
    // an equal share of the rows and none waits for a second round
    const uint64_t per_cu = std::max<uint64_t>(1, std::min<uint64_t>(2, (160 * 1024) / std::max<size_t>(lds, 1)));
    // (a directly indexed cache: twice that many, run in two waves of
    // blocks, which streams the columns faster)
    uint64_t grid = std::min<uint64_t>(ceil_div(rows, (uint64_t)kAggBlock * GE * 4),
                                       (uint64_t)cu_count() * per_cu * (a.cache_direct ? 2 : 1));
    const dim3 gd((uint32_t)std::max<uint64_t>(grid, 1));
    if (a.cache_direct && !spill && gd.x >= 64) {
      // many blocks over a cached domain: stage the caches and reduce them
      void* dsp;
      MTL_TRY(s->arena.alloc((uint64_t)gd.x * (a.cache_slots + 1) * a.W * 8, &dsp));
      a.dstage = (uint64_t*)dsp;
    }
    if (spill) {
      void *sn, *hk, *hv;
      const uint64_t hot = (uint64_t)gd.x * a.cache_slots;
      void *sk, *sp;
      MTL_TRY(s->arena.alloc(rows * 4, &sk));
      MTL_TRY(s->arena.alloc(rows * 4 * std::max<uint32_t>(a.n_pay, 1), &sp));
      a.spill_key = (uint32_t*)sk;
      a.spill_pay = (uint32_t*)sp;
      MTL_TRY(s->arena.alloc(gd.x * 4 + 4, &sn));
      MTL_TRY(s->arena.alloc(hot * 8, &hk));
      MTL_TRY(s->arena.alloc(hot * a.W * 8, &hv));
      a.spill = 1;
      a.spill_n = (uint32_t*)sn;
      a.spill_rows = rows;
      a.hot_key = (uint64_t*)hk;
      a.hot_st = (uint64_t*)hv;
    }
    {
      ProfScope ps(s, "k_group_agg");
      if (launch_agg_fast(a, gd, lds, s->stream)) {
      } else if (a.n_src <= 2) hipLaunchKernelGGL(k_group_agg<2>, gd, dim3(kAggBlock), lds, s->stream, a);
      else if (a.n_src <= 4) hipLaunchKernelGGL(k_group_agg<4>, gd, dim3(kAggBlock), lds, s->stream, a);
      else hipLaunchKernelGGL(k_group_agg<8>, gd, dim3(kAggBlock), lds, s->stream, a);
      MTL_HIP(hipGetLastError());
    }
    if (a.dstage) {
      ProfScope ps(s, "k_dense_reduce");
      const uint64_t words = (uint64_t)a.W * (a.cache_slots + 1);
      const dim3 rg((uint32_t)ceil_div(words, 256), (uint32_t)ceil_div(gd.x, kDenseReduceG));
      hipLaunchKernelGGL(k_dense_reduce, rg, dim3(256), 0, s->stream, a, dense_word_ops(a),
                         a.dstage, gd.x, a.cache_slots);
      MTL_HIP(hipGetLastError());
    }
    if (a.spill) {
      // the spilled row count sizes the partition buffers: read it with
      // the error bits
      void* hs;
      MTL_TRY(stream_pinned(s, 8 + (size_t)gd.x * 4, &hs));
      MTL_HIP(hipMemcpyAsync(hs, err, 4, hipMemcpyDeviceToHost, s->stream));
      MTL_HIP(hipMemcpyAsync((char*)hs + 8, a.spill_n, (size_t)gd.x * 4, hipMemcpyDeviceToHost,
                             s->stream));
      MTL_HIP(stream_wait(s->stream));
      uint64_t total = 0;
      for (uint32_t b = 0; b < gd.x; ++b) total += ((const uint32_t*)((char*)hs + 8))[b];
      const uint64_t seg_max = ceil_div(rows, gd.x);
      if (*(const uint32_t*)hs == 0) {
        if (a.n_src <= 2) MTL_TRY(spill_finish<2>(a, gd.x, seg_max, total, s));
        else if (a.n_src <= 4) MTL_TRY(spill_finish<4>(a, gd.x, seg_max, total, s));
        else MTL_TRY(spill_finish<8>(a, gd.x, seg_max, total, s));
      }
    }
  }
  MTL_HIP(hipGetLastError());
  g->args = a;
  g->args.err = nullptr;  // arena memory is not kept past this call
  g->args.spill = 0;
  g->args.spill_key = nullptr;
  g->args.spill_pay = nullptr;
  g->args.spill_n = nullptr;
  g->args.hot_key = nullptr;
  g->args.hot_st = nullptr;
  g->args.dstage = nullptr;
  // check errors now (synchronous boundary of the aggregate)
  void* host;
  MTL_TRY(stream_pinned(s, 8, &host));
  MTL_HIP(hipMemcpyAsync(host, err, 4, hipMemcpyDeviceToHost, s->stream));
  MTL_HIP(stream_wait(s->stream));
  uint32_t eb = *(uint32_t*)host;
  if (eb) {
    if (eb & 8u) return set_error(MTL_ERR_INVALID_ARGUMENT, "a group key lies outside its stated [min, max]");
    return set_error(MTL_ERR_PROCESSING, "processing error: group table overflow");
  }
  owner.g = nullptr;
  *out = g;
  return MTL_OK;
}

int32_t mtl_groups_release(mtl_groups* g) {
  if (!g) return MTL_OK;
  {
    // park the table's memory on its stream for the next group-by there
    mtl_stream* s = g->s;
    std::lock_guard<std::mutex> lk(s->mu);
    void* mem[2] = {g->st_mem, g->keys_mem};
    for (int i = 0; i < 2; ++i) {
      if (!mem[i]) continue;
      if (s->parked_table[i]) (void)hipFreeAsync(s->parked_table[i], s->stream);
      s->parked_table[i] = mem[i];
      s->parked_bytes[i] = g->mem_bytes[i];
    }
  }
  release_processing(g->bytes);
  delete g;
  return MTL_OK;
}

// Emit `idx` groups (device list + device count) into host buffers: n
// groups, or, with n_real set, at most n, the count read back with the rows
// in the same round trip (*n_real).
static int32_t emit_groups(mtl_groups* g, const uint32_t* idx, const uint32_t* count,
                           uint64_t n, int64_t* keys, void* const* agg_out,
                           uint32_t* n_real = nullptr, const GTopkArgs* small = nullptr) {
  mtl_stream* s = g->s;
  const GroupArgs& a = g->args;
  if (n == 0) return MTL_OK;
  // keys, counts and aggregate words in one device block, so a small result
  // comes back in one copy through the stream's pinned staging; with n_real
  // the block starts with a 256-byte header holding the device count
  const uint64_t kb = n * a.n_keys * 8, cb = n * 8, ab = n * std::max<uint32_t>(a.n_aggs, 1) * 16;
  const uint64_t hb = n_real ? 256 : 0;
  void* blob;
  MTL_TRY(s->arena.alloc(hb + kb + cb + ab, &blob));
  void* dk = (char*)blob + hb;
  void* dc = (char*)blob + hb + kb;
  void* da = (char*)blob + hb + kb + cb;
  EmitArgs e{};
  e.idx = idx;
  e.count = count;
  e.st = a.st;
  e.tkeys = a.tkeys;
  e.W = a.W;
  e.dense = a.dense;
  e.raw = a.raw;
  e.n_keys = a.n_keys;
  for (uint32_t j = 0; j < a.n_keys; ++j) {
    e.key_min[j] = a.key_min[j];
    e.key_span[j] = a.key_span[j];
    e.key_stride[j] = a.key_stride[j];
  }
  e.cap = a.cap;
  e.n_aggs = a.n_aggs;
  for (uint32_t k = 0; k < a.n_aggs; ++k) e.aggs[k] = a.aggs[k];
  e.out_keys = (int64_t*)dk;
  e.out_cnt = (uint64_t*)dc;
  e.out_aggs = (uint64_t*)da;
  e.out_n = n_real ? (uint32_t*)blob : nullptr;
  if (small) {  // the small-table top-K ranks and emits in one launch
    ProfScope ps(s, "k_group_topk_small");
    hipLaunchKernelGGL(k_group_topk_small,
                       dim3((uint32_t)ceil_div(std::max<uint64_t>(small->n, 1),
                                               (uint64_t)kSmallTopkRank)),
                       dim3(kSmallTopkThreads), 0, s->stream, *small, e);
  } else {
    ProfScope ps(s, "k_group_emit");
    hipLaunchKernelGGL(k_group_emit, dim3((uint32_t)std::min<uint64_t>(ceil_div(n, 256), 4096)),
                       dim3(256), 0, s->stream, e);
  }
  MTL_HIP(hipGetLastError());
  std::vector<uint64_t> cnt(n), st(n * std::max<uint32_t>(a.n_aggs, 1) * 2);
  constexpr uint64_t kStagedMax = 16ull << 20;
  if (kb + cb + ab <= kStagedMax) {
    void* h0;
    MTL_TRY(stream_pinned(s, 256 + kb + cb + ab, &h0));
    char* h = (char*)h0 + 256;
    MTL_HIP(hipMemcpyAsync(h - hb, blob, hb + kb + cb + ab, hipMemcpyDeviceToHost, s->stream));
    MTL_HIP(stream_wait(s->stream));
    if (n_real) {
      *n_real = *(const uint32_t*)h0;
      if (*n_real < n) n = *n_real;
    }
    if (keys) memcpy(keys, h, n * a.n_keys * 8);
    memcpy(cnt.data(), h + kb, n * 8);
    memcpy(st.data(), h + kb + cb, n * std::max<uint32_t>(a.n_aggs, 1) * 16);
  } else {
    if (n_real) return set_error(MTL_ERR_INVALID_ARGUMENT, "emission of %llu groups",
                                 (unsigned long long)n);
    if (keys) MTL_HIP(hipMemcpyAsync(keys, dk, kb, hipMemcpyDeviceToHost, s->stream));
    MTL_HIP(hipMemcpyAsync(cnt.data(), dc, cb, hipMemcpyDeviceToHost, s->stream));
    MTL_HIP(hipMemcpyAsync(st.data(), da, ab, hipMemcpyDeviceToHost, s->stream));
    MTL_HIP(stream_wait(s->stream));
  }
  if (!agg_out) return MTL_OK;
  for (uint32_t k = 0; k < a.n_aggs; ++k) {
    if (!agg_out[k]) continue;
    const AggSlot& sl = a.aggs[k];
    int32_t func = g->agg_func[k];
    for (uint64_t i = 0; i < n; ++i) {
      const uint64_t* v = &st[(i * a.n_aggs + k) * 2];
      switch (func) {
        case MTL_AGG_COUNT:
          ((int64_t*)agg_out[k])[i] = (int64_t)cnt[i];
          break;
        case MTL_AGG_SUM:
          if (sl.kind == K_SUMI) {
            ((uint64_t*)agg_out[k])[2 * i] = v[0];
            ((uint64_t*)agg_out[k])[2 * i + 1] = v[1];
          } else {
            double d;
            memcpy(&d, &v[0], 8);
            ((double*)agg_out[k])[i] = d;
          }
          break;
        case MTL_AGG_AVG:
          if (sl.kind == K_SUMI) {
            // hugeint sum / (count * 10^scale) in long double, as DuckDB's
            // integer and DECIMAL AVG
            __int128 sum = (__int128)(((unsigned __int128)v[1] << 64) | v[0]);
            ((double*)agg_out[k])[i] = (double)((long double)sum /
                                                ((long double)cnt[i] * pow10_ld(g->agg_scale[k])));
          } else {
            double d;
            memcpy(&d, &v[0], 8);
            ((double*)agg_out[k])[i] = d / (double)cnt[i];
          }
          break;
        default:  // MIN / MAX
          if (sl.kind == K_MINF || sl.kind == K_MAXF) {
            uint64_t b = v[0];
            b = (b & 0x8000000000000000ull) ? (b & ~0x8000000000000000ull) : ~b;
            double d;
            memcpy(&d, &b, 8);
            ((double*)agg_out[k])[i] = d;
          } else {
            ((int64_t*)agg_out[k])[i] = (int64_t)v[0];
          }
          break;
      }
    }
  }
  return MTL_OK;
}

static int32_t compact_groups(mtl_groups* g, uint32_t** idx, uint32_t** count) {
  mtl_stream* s = g->s;
  const GroupArgs& a = g->args;
  uint64_t n = g->entries;
  uint32_t tiles = (uint32_t)ceil_div(n ? n : 1, kTileRows);
  void *list, *bm, *counts, *misc;
  MTL_TRY(s->arena.alloc((n ? n : 1) * 4, &list));
  MTL_TRY(s->arena.alloc((uint64_t)tiles * kTileRows / 8, &bm));
  MTL_TRY(s->arena.alloc((uint64_t)tiles * 4, &counts));
  MTL_TRY(s->arena.alloc(256, &misc));
  {
    ProfScope ps(s, "k_group_compact");
    hipLaunchKernelGGL(k_group_flags, dim3(tiles), dim3(256), 0, s->stream, a.st, a.W, n,
                       (uint64_t*)bm, (uint32_t*)counts);
    MTL_HIP(hipGetLastError());
    MTL_TRY(launch_compact((const uint64_t*)bm, (uint32_t*)counts, tiles, nullptr,
                           (uint32_t*)list, (uint32_t*)misc, s->stream));
  }
  *idx = (uint32_t*)list;
  *count = (uint32_t*)misc;
  return MTL_OK;
}

int32_t mtl_groups_count(mtl_groups* g, uint64_t* n) {
  if (!g || !n) return set_error(MTL_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(g->s->mu);
  g->s->arena.reset();
  uint32_t *idx, *count;
  MTL_TRY(compact_groups(g, &idx, &count));
  void* host;
  MTL_TRY(stream_pinned(g->s, 8, &host));
  MTL_HIP(hipMemcpyAsync(host, count, 4, hipMemcpyDeviceToHost, g->s->stream));
  MTL_HIP(stream_wait(g->s->stream));
  *n = *(uint32_t*)host;
  return MTL_OK;
}

int32_t mtl_groups_fetch(mtl_groups* g, int64_t* keys, void* const* agg_out, uint64_t capacity,
                         uint64_t* n_out) {
  if (!g || !n_out) return set_error(MTL_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(g->s->mu);
  g->s->arena.reset();
  uint32_t *idx, *count;
  MTL_TRY(compact_groups(g, &idx, &count));
  void* host;
  MTL_TRY(stream_pinned(g->s, 8, &host));
  MTL_HIP(hipMemcpyAsync(host, count, 4, hipMemcpyDeviceToHost, g->s->stream));
  MTL_HIP(stream_wait(g->s->stream));
  uint64_t n = *(uint32_t*)host;
  *n_out = n;
  if (n > capacity)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "%llu groups exceed the capacity %llu",
                     (unsigned long long)n, (unsigned long long)capacity);
  return emit_groups(g, idx, count, n, keys, agg_out);
}

int32_t mtl_groups_topk(mtl_groups* g, uint32_t agg, int32_t desc, uint32_t k, int64_t* keys,
                        void* const* agg_out, uint32_t* n_out) {
  if (!g || !n_out) return set_error(MTL_ERR_INVALID_ARGUMENT, "NULL argument");
  const GroupArgs& a = g->args;
  if (agg >= a.n_aggs) return set_error(MTL_ERR_INVALID_ARGUMENT, "aggregate index %u", agg);
  if (k == 0 || k > (uint32_t)TOPK_MAX_K)
    return set_error(MTL_UNSUPPORTED, "top-K with K=%u (supported 1..%d)", k, TOPK_MAX_K);
  mtl_stream* s = g->s;
  std::lock_guard<std::mutex> lk(s->mu);
  s->arena.reset();
  void *misc, *oidx;
  MTL_TRY(s->arena.alloc(256, &misc));
  MTL_TRY(s->arena.alloc((uint64_t)k * 4, &oidx));
  GTopkArgs t{};
  t.st = a.st;
  t.tkeys = a.tkeys;
  t.W = a.W;
  t.dense = a.dense;
  t.raw = a.raw;
  t.n = g->entries;
  t.cap = a.cap;
  t.kind = a.aggs[agg].kind;
  t.word = a.aggs[agg].word;
  t.desc = desc;
  t.avg = g->agg_func[agg] == MTL_AGG_AVG;
  t.avg_scale = 1;
  if (t.avg && t.kind == K_SUMI) {
    // the divisor count * 10^scale must stay an exact integer below 2^64
    // (count < 2^32): DECIMAL scales up to 9
    if (g->agg_scale[agg] < 0 || g->agg_scale[agg] > 9)
      return set_error(MTL_UNSUPPORTED, "ORDER BY AVG of a DECIMAL with scale %d",
                       g->agg_scale[agg]);
    for (int i = 0; i < g->agg_scale[agg]; ++i) t.avg_scale *= 10;
  }
  t.k = k;
  t.out_idx = (uint32_t*)oidx;
  t.out_n = (uint32_t*)misc + 1;
  if (g->entries <= kSmallTopk)
    return emit_groups(g, nullptr, nullptr, k, keys, agg_out, n_out, &t);
  const uint64_t nt = k <= 128 ? 1024 : 256;
  uint32_t grid = (uint32_t)std::max<uint64_t>(
      1, std::min<uint64_t>(ceil_div(g->entries, nt * 16), (uint64_t)cu_count()));
  void *ck[4], *cid, *lk0[2], *lg[2], *ln[2];
  const uint64_t stage_n = (uint64_t)grid * (nt / 64) * k;
  for (int f = 0; f < 4; ++f) MTL_TRY(s->arena.alloc(stage_n * 8, &ck[f]));
  MTL_TRY(s->arena.alloc(stage_n * 4, &cid));
  const uint64_t la = grid, lb = ceil_div(grid, topk_fan(k));
  MTL_TRY(s->arena.alloc(la * k * 8, &lk0[0]));
  MTL_TRY(s->arena.alloc(la * k * 4, &lg[0]));
  MTL_TRY(s->arena.alloc(la * 4, &ln[0]));
  MTL_TRY(s->arena.alloc(lb * k * 8, &lk0[1]));
  MTL_TRY(s->arena.alloc(lb * k * 4, &lg[1]));
  MTL_TRY(s->arena.alloc(lb * 4, &ln[1]));
  ZeroList zl;
  zl.add((uint32_t*)misc, 64);
  MTL_TRY(launch_zero(zl, s->stream));
  for (int f = 0; f < 4; ++f) t.stage.ck[f] = (uint64_t*)ck[f];
  t.stage.cid = (uint32_t*)cid;
  TopkLists lists[2];
  for (int b = 0; b < 2; ++b)
    lists[b] = TopkLists{(uint64_t*)lk0[b], (uint32_t*)lg[b], (uint32_t*)ln[b]};
  t.lists0 = lists[0];
  {
    ProfScope ps(s, "k_group_topk");
    if (k <= 64)
      hipLaunchKernelGGL(k_group_topk<1>, dim3(grid), dim3(TopkShape<1>::NT), 0, s->stream, t);
    else if (k <= 128)
      hipLaunchKernelGGL(k_group_topk<2>, dim3(grid), dim3(TopkShape<2>::NT), 0, s->stream, t);
    else
      hipLaunchKernelGGL(k_group_topk<8>, dim3(grid), dim3(TopkShape<8>::NT), 0, s->stream, t);
  }
  MTL_HIP(hipGetLastError());
  {
    ProfScope ps(s, "k_group_topk_merge");
    const uint32_t fan = topk_fan(k);
    uint32_t L = grid;
    int in = 0;
    while (true) {
      const uint32_t blocks = (uint32_t)ceil_div(L, fan);
      const uint32_t fin = blocks == 1 ? 1u : 0u;
      hipLaunchKernelGGL(k_group_topk_merge, dim3(blocks), dim3(TOPK_MERGE_THREADS), 0,
                         s->stream, t, lists[in], lists[in ^ 1], L, fin);
      MTL_HIP(hipGetLastError());
      if (fin) break;
      L = blocks;
      in ^= 1;
    }
  }
  // the rows of all k candidates come back with the count in one round trip
  return emit_groups(g, (const uint32_t*)oidx, (uint32_t*)misc + 1, k, keys, agg_out, n_out);
}

}  // extern "C"
