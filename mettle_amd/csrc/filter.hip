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
#include "rowtile.hpp"

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
// Filter + compaction over row tiles (rowtile.hpp).
// ---------------------------------------------------------------------------
// Evaluate all predicates (and the optional semi-join probe) for the 16 rows
// of this lane; returns the pass bits (bit 4j+v).
template <bool SEL>
__device__ __forceinline__ uint32_t eval_lane(const FilterArgs& a, uint64_t n,
                                              uint64_t k0,
                                              uint32_t (&rid)[FJ][4]) {
  uint32_t bits = 0;
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (k0 + (uint64_t)j * 256 + v < n) bits |= 1u << (4 * j + v);
  if (SEL) {
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
    // rows past the end read row 0 (always in bounds); their bits are clear
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        uint64_t k = k0 + (uint64_t)j * 256 + v;
        rid[j][v] = k < n ? (uint32_t)k : 0u;
      }
  }

  for (int p = 0; p < a.n_preds && bits; ++p) {
    const bool vec_ok = (a.vec_ok >> p) & 1u;
    if (SEL) pred_sel(a.preds[p], rid, bits);
    else pred_dense(a.preds[p], k0, n, vec_ok, bits);
  }
  // Semi-join probe: key(row) - min must be a set bit of the bitmap.  All
  // 16 keys and bitmap words are loaded before any is used, so the loads
  // overlap instead of serialising behind one another.
  if (a.probe_bm && bits) {
    int64_t key[FJ][4];
    load_keys16(a.probe_key, a.probe_key_type, !SEL && a.probe_vec, k0, n, rid, key);
    uint64_t w[FJ][4];
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        uint64_t s = key_off(key[j][v], a.probe_min);
        w[j][v] = a.probe_bm[s < a.probe_span ? (s >> 6) : 0];
      }
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        uint64_t s = key_off(key[j][v], a.probe_min);
        bool hit = s < a.probe_span && ((w[j][v] >> (s & 63)) & 1ull);
        if (!hit) bits &= ~(1u << (4 * j + v));
      }
  }
  return bits;
}

// Words of the mark bitmap near the tile's first key are OR-ed in an LDS
// window and flushed once per block (sorted keys: one global atomic per
// touched word instead of one per lane-step); others go straight to HBM.
constexpr int kMarkWin = 1024;  // words (64K keys), 8 KB of LDS

__device__ __forceinline__ void mark_word(const FilterArgs& a, uint64_t* win, uint64_t wbase,
                                          uint64_t word, uint64_t bits) {
  uint64_t wi = word - wbase;
  if (wi < (uint64_t)kMarkWin) atomicOr((unsigned long long*)&win[wi], bits);
  else atomicOr((unsigned long long*)&a.mark_bm[word], bits);
}

// Set bit (key - min) of the mark bitmap for every kept row.  The (up to)
// four rows of a lane-step usually share one 64-bit word (sorted keys):
// their bits are OR-ed first, so one atomic covers them.  The atomics do
// not return: a repeated key is detected afterwards by comparing the
// bitmap's popcount with the popcount of the pass bits (sel_bm).
// ---------------------------------------------------------------------------
// Narrow dense path (no input selection; at most one predicate; its column
// and the probe and mark keys are at most 4 bytes wide, key spans < 2^32): the
// three columns are loaded together, one round trip, into 32-bit registers,
// and the probe bitmap is read as 32-bit words, so the kernel keeps enough
// registers free for a high occupancy.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t eval_narrow(const FilterArgs& a, uint64_t n, uint64_t k0,
                                                uint32_t (&ms)[FJ][4]) {
  uint32_t bits = 0;
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (k0 + (uint64_t)j * 256 + v < n) bits |= 1u << (4 * j + v);
  uint32_t px[FJ][4], pk[FJ][4];
  if (a.n_preds)
    load_u32x16(a.preds[0].col, pred_width(a.preds[0].kind), k0, n, a.vec_ok & 1u, px);
  if (a.probe_bm)
    load_u32x16(a.probe_key, a.probe_key_type == MTL_DICT8 ? 1u : 4u, k0, n, a.probe_vec, pk);
  if (a.mark_bm)
    load_u32x16(a.mark_key, a.mark_key_type == MTL_DICT8 ? 1u : 4u, k0, n, a.mark_vec, ms);
  // the predicate first (its column arrived with the keys), so the bitmap
  // probe below reads only rows that still pass: random probes are L2
  // line reads, the largest cost of this pass
  if (a.n_preds) {
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (!pred_pass(enc_raw(px[j][v], a.preds[0].kind), a.preds[0]))
          bits &= ~(1u << (4 * j + v));
  }
  if (a.probe_bm) {
    const uint32_t* bm32 = (const uint32_t*)a.probe_bm;
    uint32_t w[FJ][4];
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        pk[j][v] = narrow_off(pk[j][v], a.probe_key_type, a.probe_min, a.probe_span);
        const bool need = ((bits >> (4 * j + v)) & 1u) && pk[j][v] != 0xFFFFFFFFu;
        w[j][v] = need ? ld_off32(bm32, (pk[j][v] >> 5) << 2) : 0u;
      }
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (pk[j][v] == 0xFFFFFFFFu || !((w[j][v] >> (pk[j][v] & 31)) & 1u))
          bits &= ~(1u << (4 * j + v));
  }
  if (a.mark_bm) {
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        ms[j][v] = narrow_off(ms[j][v], a.mark_key_type, a.mark_min, a.mark_span);
  }
  return bits;
}

// Mark the offsets of the kept rows (0xFFFFFFFF: key outside the range).
__device__ __forceinline__ void mark_offsets(const FilterArgs& a, const uint32_t (&ms)[FJ][4],
                                             uint32_t bits, uint64_t* win, uint64_t wbase) {
  uint32_t e = 0;
#pragma unroll
  for (int j = 0; j < FJ; ++j) {
    uint32_t m = (bits >> (4 * j)) & 0xFu;
    if (!m) continue;
    uint64_t wj = ~0ull, bj = 0;
    uint32_t rest = 0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      if (!((m >> v) & 1u)) continue;
      const uint32_t s = ms[j][v];
      if (s == 0xFFFFFFFFu) {
        e |= 8u;
        continue;
      }
      if (wj == ~0ull) wj = s >> 6;
      if ((s >> 6) == wj) bj |= 1ull << (s & 63);
      else rest |= 1u << v;
    }
    if (bj) mark_word(a, win, wbase, wj, bj);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      if (!((rest >> v) & 1u)) continue;
      const uint32_t s = ms[j][v];
      mark_word(a, win, wbase, s >> 6, 1ull << (s & 63));
    }
  }
  if (e) atomicOr(a.err, e);
}

template <bool SEL>
__device__ __forceinline__ void mark_rows(const FilterArgs& a, uint64_t n, uint64_t k0,
                                          const uint32_t (&rid)[FJ][4], uint32_t bits,
                                          uint64_t* win, uint64_t wbase) {
  int64_t key[FJ][4];
  load_keys16(a.mark_key, a.mark_key_type, !SEL && a.mark_vec, k0, n, rid, key);
  uint32_t e = 0;
#pragma unroll
  for (int j = 0; j < FJ; ++j) {
    uint32_t m = (bits >> (4 * j)) & 0xFu;
    if (!m) continue;
    uint64_t wj = ~0ull, bj = 0;
    uint32_t rest = 0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      if (!((m >> v) & 1u)) continue;
      uint64_t s = key_off(key[j][v], a.mark_min);
      if (s >= a.mark_span) {
        e |= 8u;
        continue;
      }
      if (wj == ~0ull) wj = s >> 6;
      if ((s >> 6) == wj) bj |= 1ull << (s & 63);
      else rest |= 1u << v;
    }
    if (bj) mark_word(a, win, wbase, wj, bj);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      if (!((rest >> v) & 1u)) continue;
      uint64_t s = key_off(key[j][v], a.mark_min);
      mark_word(a, win, wbase, s >> 6, 1ull << (s & 63));
    }
  }
  if (e) atomicOr(a.err, e);
}

// Pass 1 of a filter: one tile of 4096 rows per block, no inter-block
// dependency.  Writes the tile's pass bits to sel_bm (1 bit per row, bit r
// for row r of the input order) and its count to tile_counts[tile].
template <bool SEL, bool NARROW>
__global__ __launch_bounds__(FBLOCK) void k_filter(FilterArgs a) {
  __shared__ uint32_t s_wave_tot[FWAVES];
  const int tid = threadIdx.x;
  const int wave = tid / WAVE;
  const int lane = tid % WAVE;
  const uint32_t tile = blockIdx.x;
  const uint64_t n = SEL ? dev_count(a.in_count, a.n_rows) : a.n_rows;
  const uint64_t seg = (uint64_t)tile * FTILE + (uint64_t)wave * (FJ * 256);
  const uint64_t k0 = seg + (uint64_t)lane * 4;
  uint32_t rid[FJ][4];  // row ids (generic path) or mark offsets (narrow path)
  uint32_t bits = 0;
  if ((uint64_t)tile * FTILE < n) {
    if (NARROW) bits = eval_narrow(a, n, k0, rid);
    else bits = eval_lane<SEL>(a, n, k0, rid);
  } else {
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) rid[j][v] = NARROW ? 0xFFFFFFFFu : 0u;
  }
  if (a.mark_bm) {
    __shared__ uint64_t win[kMarkWin];
    __shared__ uint64_t s_wbase;
    for (int i = tid; i < kMarkWin; i += FBLOCK) win[i] = 0;
    if (NARROW) {
      // the window starts at the word of the tile's first row (tid 0's
      // first mark offset, already in registers)
      if (tid == 0) s_wbase = rid[0][0] != 0xFFFFFFFFu ? (rid[0][0] >> 6) : 0;
    } else if (tid == 0) {
      uint64_t base = 0;
      if ((uint64_t)tile * FTILE < n) {
        uint64_t r = SEL ? a.in_sel[(uint64_t)tile * FTILE] : (uint64_t)tile * FTILE;
        uint64_t s0 = key_off(load_key(a.mark_key, a.mark_key_type, r), a.mark_min);
        base = s0 < a.mark_span ? (s0 >> 6) : 0;
      }
      s_wbase = base;
    }
    __syncthreads();
    const uint64_t wbase = s_wbase;
    if (NARROW) mark_offsets(a, rid, bits, win, wbase);
    else mark_rows<SEL>(a, n, k0, rid, bits, win, wbase);
    __syncthreads();
    const uint64_t words = (a.mark_span + 63) / 64;
    for (int i = tid; i < kMarkWin; i += FBLOCK) {
      uint64_t w = win[i];
      if (w && wbase + i < words) atomicOr((unsigned long long*)&a.mark_bm[wbase + i], w);
    }
  }
  if (a.sel_bm) store_pass_bits(a.sel_bm, seg, bits, lane);
  if (a.tile_counts) {
    uint32_t c = wave_sum_u32((uint32_t)__popc(bits));
    if (lane == 0) s_wave_tot[wave] = c;
    __syncthreads();
    if (tid == 0) {
      uint32_t t = 0;
#pragma unroll
      for (int w = 0; w < FWAVES; ++w) t += s_wave_tot[w];
      a.tile_counts[tile] = t;
      if (a.group_sums && t) atomicAdd(&a.group_sums[(tile >> kGroupShift) * kGroupStride], t);
    }
  }
}

// ---------------------------------------------------------------------------
// Persistent probe pass.  A narrow pass with a semi-join probe and no tile
// counts (the parent pass of a star join) is held back by its probes: a
// probe of a random key is a divergent load, one L2 request per lane, and
// probing about half the rows of a 15M-row parent took 19 of the tiled
// pass's 68 us.  Here one block of 1024 threads per CU copies the low words
// of the probe bitmap (up to 144 KB: 1,179,648 keys) into LDS once; its waves
// then walk 1024-row segments round-robin, and probes of keys in the copied
// part read LDS.  Marks are OR-combined across the lanes of each 16-lane row
// that hit the same 64-bit word (a DPP segmented OR scan), so a run of
// sorted keys costs one atomic per word, and no barrier follows the copy.
// ---------------------------------------------------------------------------
constexpr int PBLOCK = 1024;
constexpr uint32_t kLdsProbeWords = 36864;  // 32-bit words, 144 KB

// Mark the offsets of the kept rows of one wave segment.  Per step, each
// lane folds its rows that share the word of its first kept row into one
// (word, bits) pair; a segmented OR scan over each 16-lane row (DPP row
// shifts) merges the pairs of lanes with equal words (OR is idempotent, so
// merging equal words that are not adjacent is harmless), and the last lane
// of each run of equal words in a row issues the atomic.  With dense sorted
// keys a row's 64 rows are exactly one bitmap word: one atomic per word.  A
// lane's rows with another word are marked one by one.
// One step of the segmented OR scan: take the bits of lane r - OFF of the
// row when it holds the same word.
template <int OFF>
__device__ __forceinline__ void row_seg_or(uint32_t r, uint32_t w, uint32_t& lo, uint32_t& hi) {
  const uint32_t ow = row_shr<OFF>(w), ol = row_shr<OFF>(lo), oh = row_shr<OFF>(hi);
  if (r >= (uint32_t)OFF && ow == w) {
    lo |= ol;
    hi |= oh;
  }
}

__device__ __forceinline__ void mark_offsets_wave(const FilterArgs& a,
                                                  const uint32_t (&ms)[FJ][4], uint32_t bits,
                                                  int lane) {
  uint32_t e = 0;
#pragma unroll
  for (int j = 0; j < FJ; ++j) {
    const uint32_t m = (bits >> (4 * j)) & 0xFu;
    if (__ballot(m != 0) == 0) continue;  // wave-uniform: nothing kept at this step
    // branch-free fold: the word of the first kept row (selects from the
    // last row back), then the bits of the kept rows in that word
    uint32_t wj = 0xFFFFFFFFu, rest = 0, lo = 0, hi = 0;
#pragma unroll
    for (int v = 3; v >= 0; --v) {
      const bool k = ((m >> v) & 1u) && ms[j][v] != 0xFFFFFFFFu;
      wj = k ? ms[j][v] >> 6 : wj;
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const uint32_t sv = ms[j][v];
      const bool sel = (m >> v) & 1u;
      e |= (sel && sv == 0xFFFFFFFFu) ? 8u : 0u;
      const bool k = sel && sv != 0xFFFFFFFFu;
      const bool same = k && (sv >> 6) == wj;
      const uint32_t b = 1u << (sv & 31);
      lo |= (same && !(sv & 32)) ? b : 0u;
      hi |= (same && (sv & 32)) ? b : 0u;
      rest |= (k && !same) ? 1u << v : 0u;
    }
    const uint32_t r = (uint32_t)lane & 15u;
    row_seg_or<1>(r, wj, lo, hi);
    row_seg_or<2>(r, wj, lo, hi);
    row_seg_or<4>(r, wj, lo, hi);
    row_seg_or<8>(r, wj, lo, hi);
    const uint32_t nw = row_shl1(wj);
    if ((lo | hi) && (r == 15 || nw != wj))
      atomicOr((unsigned long long*)&a.mark_bm[wj], ((uint64_t)hi << 32) | lo);
    if (__ballot(rest != 0) == 0) continue;  // sorted keys: a lane's rows share a word
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      if (!((rest >> v) & 1u)) continue;
      const uint32_t sv = ms[j][v];
      atomicOr((unsigned long long*)&a.mark_bm[sv >> 6], 1ull << (sv & 63));
    }
  }
  if (e) atomicOr(a.err, e);
}

// The 16 values of lane l of a 1024-row wave segment of a 4-byte column
// (rows k0 + 256 j + v, k0 = seg + 4 l).  FULL: the segment lies inside the
// table and the column is 16-byte aligned, so each step is one streaming
// dwordx4 with no branch, and the loads of all columns go out back to back.
template <bool FULL>
__device__ __forceinline__ void seg_load4(const uint32_t* col, uint64_t k0, uint64_t n,
                                          uint32_t (&x)[FJ][4]) {
#pragma unroll
  for (int j = 0; j < FJ; ++j) {
    const uint64_t k = k0 + (uint64_t)j * 256;
    if (FULL) {
      const uint4 u = ld_stream(col + k);
      x[j][0] = u.x; x[j][1] = u.y; x[j][2] = u.z; x[j][3] = u.w;
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v) x[j][v] = k + v < n ? col[k + v] : 0u;
    }
  }
}

// Offset of a 4-byte integer key from min, 0xFFFFFFFF outside [0, span).
// K32 (the host checked 0 <= min and min + span <= 2^31): one 32-bit
// subtraction, since a negative key or one below min wraps to at least
// 2^31 - min >= span.
template <bool K32>
__device__ __forceinline__ uint32_t key_off4(uint32_t raw, int64_t min, uint64_t span) {
  if (K32) {
    const uint32_t o = raw - (uint32_t)min;
    return o < (uint32_t)span ? o : 0xFFFFFFFFu;
  }
  return narrow_off(raw, MTL_INTEGER, min, span);
}

// One wave segment of the probe pass: returns the pass bits (bit 4j+v) and
// leaves the mark offsets in ms (0xFFFFFFFF: key outside the range).
// Predicate (of kind PK; none when PK < 0) first, then the probe of the
// surviving rows: offsets
// below lds_keys read the LDS copy, the rest the bitmap in HBM.  Every probe
// load is issued unconditionally (a lane with nothing to probe reads word 0)
// and the results are selected afterwards, so no branch separates the loads.
template <int PK, bool K32, bool FULL>
__device__ __forceinline__ uint32_t probe_segment(const FilterArgs& a, uint64_t n, uint64_t k0,
                                                  const uint32_t* lds, uint32_t lds_keys,
                                                  uint32_t (&ms)[FJ][4]) {
  uint32_t bits = 0;
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (FULL || k0 + (uint64_t)j * 256 + v < n) bits |= 1u << (4 * j + v);
  uint32_t px[FJ][4], pk[FJ][4];
  if (PK >= 0) seg_load4<FULL>((const uint32_t*)a.preds[0].col, k0, n, px);
  seg_load4<FULL>((const uint32_t*)a.probe_key, k0, n, pk);
  seg_load4<FULL>((const uint32_t*)a.mark_key, k0, n, ms);
  // keep the scheduler from trading the loads of later steps for registers:
  // all twelve go out before any of them is used
  __builtin_amdgcn_sched_barrier(0);
  if (PK >= 0) {
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (!pred_pass(enc_raw(px[j][v], (uint32_t)PK), a.preds[0]))
          bits &= ~(1u << (4 * j + v));
  }
  const uint32_t* bm32 = (const uint32_t*)a.probe_bm;
  uint32_t wg[FJ][4];
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const uint32_t o = key_off4<K32>(pk[j][v], a.probe_min, a.probe_span);
      pk[j][v] = ((bits >> (4 * j + v)) & 1u) ? o : 0xFFFFFFFFu;
      const bool in_hbm = pk[j][v] >= lds_keys && pk[j][v] != 0xFFFFFFFFu;
      wg[j][v] = ld_off32(bm32, in_hbm ? (pk[j][v] >> 5) << 2 : 0u);
    }
  __builtin_amdgcn_sched_barrier(0);
  // the LDS reads go out while the HBM probes are in flight
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const uint32_t o = pk[j][v];
      const uint32_t wl = lds[o < lds_keys ? o >> 5 : 0u];
      const uint32_t w = o < lds_keys ? wl : wg[j][v];
      const uint32_t hit = (uint32_t)(o != 0xFFFFFFFFu) & (w >> (o & 31));
      bits &= ~((~hit & 1u) << (4 * j + v));
      ms[j][v] = key_off4<K32>(ms[j][v], a.mark_min, a.mark_span);
    }
  return bits;
}

// lds_words: 32-bit words of the probe bitmap to copy (a multiple of 4, at
// most kLdsProbeWords and at most the bitmap's size).  The host sends only
// passes whose probe and mark keys are 4-byte integers, with one predicate
// of kind PK (PK_I32 or PK_F32) or none (PK < 0), and a mark bitmap.
template <int PK, bool K32>
__global__ __launch_bounds__(PBLOCK) void k_filter_probe(FilterArgs a, uint32_t lds_words) {
  __shared__ uint32_t s_bm[kLdsProbeWords];
  const int tid = threadIdx.x, wave = tid / WAVE, lane = tid % WAVE;
  const uint32_t* bm32 = (const uint32_t*)a.probe_bm;
  {
    // all nine 16-byte loads of a thread in flight together
    constexpr int kCopy = kLdsProbeWords / (PBLOCK * 4);
    static_assert(kCopy * PBLOCK * 4 == kLdsProbeWords, "copy covers the LDS words");
    uint4 t[kCopy];
#pragma unroll
    for (int c = 0; c < kCopy; ++c) {
      const uint32_t i = ((uint32_t)c * PBLOCK + (uint32_t)tid) * 4;
      t[c] = *(const uint4*)&bm32[i < lds_words ? i : 0u];  // unconditional: no branch
    }
#pragma unroll
    for (int c = 0; c < kCopy; ++c) {
      const uint32_t i = ((uint32_t)c * PBLOCK + (uint32_t)tid) * 4;
      if (i < lds_words) *(uint4*)&s_bm[i] = t[c];
    }
  }
  __syncthreads();
  const uint32_t lds_keys = lds_words * 32;
  const uint64_t n = a.n_rows;
  const bool vec = a.probe_vec && a.mark_vec && (PK < 0 || (a.vec_ok & 1u));
  // every segment of the 4096-row tiles, so the pass bits of rows past n up
  // to the tile end are written (as zeros) just as the tiled kernel does
  const uint64_t segs = (n + FTILE - 1) / FTILE * (FTILE / (FJ * 256));
  const uint64_t nwaves = (uint64_t)gridDim.x * (PBLOCK / WAVE);
  for (uint64_t sg = (uint64_t)blockIdx.x * (PBLOCK / WAVE) + wave; sg < segs; sg += nwaves) {
    const uint64_t seg = sg * (FJ * 256);
    const uint64_t k0 = seg + (uint64_t)lane * 4;
    uint32_t ms[FJ][4];
    uint32_t bits;
    if (vec && seg + FJ * 256 <= n)
      bits = probe_segment<PK, K32, true>(a, n, k0, s_bm, lds_keys, ms);
    else bits = probe_segment<PK, K32, false>(a, n, k0, s_bm, lds_keys, ms);
    mark_offsets_wave(a, ms, bits, lane);
    store_pass_bits(a.sel_bm, seg, bits, lane);
  }
}

// Pass 2: exclusive scan of the tile counts in place, total to *total.  One
// block of 1024: each thread owns a run of SRUN consecutive counts (serial
// sums), then one block-wide scan of the thread totals per round of
// 1024 * SRUN counts, with a running carry between rounds.
constexpr int SBLOCK = 1024;
constexpr int SRUN = 16;
__global__ __launch_bounds__(SBLOCK) void k_scan_counts(uint32_t* counts, uint32_t T,
                                                        uint32_t* total) {
  __shared__ uint32_t s_wave[SBLOCK / WAVE];
  const int tid = threadIdx.x, wave = tid / WAVE, lane = tid % WAVE;
  uint32_t carry = 0;
  for (uint32_t g = 0; g < T; g += SBLOCK * SRUN) {
    const uint32_t i0 = g + (uint32_t)tid * SRUN;
    uint32_t v[SRUN];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < SRUN; ++k) {
      v[k] = i0 + k < T ? counts[i0 + k] : 0u;
      sum += v[k];
    }
    const uint32_t incl = wave_inclusive_u32(sum);
    if (lane == WAVE - 1) s_wave[wave] = incl;
    __syncthreads();
    uint32_t off = 0, all = 0;
#pragma unroll
    for (int w = 0; w < SBLOCK / WAVE; ++w) {
      const uint32_t t = s_wave[w];
      off += w < wave ? t : 0u;
      all += t;
    }
    uint32_t run = carry + off + incl - sum;
#pragma unroll
    for (int k = 0; k < SRUN; ++k) {
      if (i0 + k < T) counts[i0 + k] = run;
      run += v[k];
    }
    carry += all;
    __syncthreads();
  }
  if (tid == 0) *total = carry;
}

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

// Optional projection fused into the expansion: out[i] = col[row_i] op s.
struct ExpandProj {
  const void* col;
  int32_t op;
  int64_t scalar;  // int64, or double bits for FLOAT/DOUBLE
  void* out;
  uint32_t* ovf;
};

// Fused projection of row r: the column value (raw bits), then the result
// written at output position o.  PK: 0 none, 1 int32, 2 int64, 3 float,
// 4 double.
template <int PK>
__device__ __forceinline__ uint64_t proj_load(const ExpandProj& pj, uint32_t r) {
  if constexpr (PK == 1 || PK == 3) return ((const uint32_t*)pj.col)[r];
  else if constexpr (PK == 2 || PK == 4) return ((const uint64_t*)pj.col)[r];
  else return 0;
}
// Returns the stored value's bits (zero-extended), or 0 without projection.
template <int PK>
__device__ __forceinline__ uint64_t proj_store(const ExpandProj& pj, uint64_t o, uint64_t x,
                                               bool& ovf) {
  if constexpr (PK == 1) {
    const int32_t y = proj_int<int32_t>((int32_t)(uint32_t)x, (int32_t)pj.scalar, pj.op, ovf);
    ((int32_t*)pj.out)[o] = y;
    return (uint32_t)y;
  } else if constexpr (PK == 2) {
    const int64_t y = proj_int<int64_t>((int64_t)x, pj.scalar, pj.op, ovf);
    ((int64_t*)pj.out)[o] = y;
    return (uint64_t)y;
  } else if constexpr (PK == 3) {
    const float y = proj_fp<float>(__builtin_bit_cast(float, (uint32_t)x),
                                   (float)__longlong_as_double(pj.scalar), pj.op);
    ((float*)pj.out)[o] = y;
    return __builtin_bit_cast(uint32_t, y);
  } else if constexpr (PK == 4) {
    const double y =
        proj_fp<double>(__builtin_bit_cast(double, x), __longlong_as_double(pj.scalar), pj.op);
    ((double*)pj.out)[o] = y;
    return __builtin_bit_cast(uint64_t, y);
  } else {
    return 0;
  }
}

// Tiles are grouped 256 to a group: the filter pass adds each tile's count
// to its group's sum (one non-returning atomic per tile, 256 per address,
// one cache line per group),
// so a tile's output offset is the sum of the groups before its own plus
// the counts of the tiles before it in its group, at most 255 + tiles/256
// loads spread over one block.  This replaces a scan pass (or a decoupled
// look-back, whose depth grows with the thousands of tiles in flight).

__device__ __forceinline__ uint32_t tile_offset(const uint32_t* __restrict__ counts,
                                                const uint32_t* __restrict__ group_sums,
                                                uint32_t tile, uint32_t* s_red) {
  const int tid = threadIdx.x, wave = tid / WAVE, lane = tid % WAVE;
  const uint32_t g = tile >> kGroupShift, first = g << kGroupShift;
  uint32_t part = first + tid < tile ? counts[first + tid] : 0u;
  for (uint32_t i = tid; i < g; i += FBLOCK) part += group_sums[i * kGroupStride];
  part = wave_sum_u32(part);
  if (lane == 0) s_red[wave] = part;
  __syncthreads();
  uint32_t off = 0;
#pragma unroll
  for (int w = 0; w < FWAVES; ++w) off += s_red[w];
  return off;
}

// One 4096-row tile per block: the positions of the tile's pass bits are a
// block prefix sum; the row ids go to LDS first, then to out with one
// coalesced store per thread and position (consecutive lanes, consecutive
// addresses), and the projection, when fused, gathers its column in row
// order in the same loop, four positions per thread in flight.  The tile's
// output offset is offsets[tile] (scanned counts) or, with group_sums, is
// computed here from the raw counts; then the last tile writes *out_count.
// With FP the tile adds the fingerprint terms of its positions (the value
// is the projection, or the row id without one) to its group's sum.
template <int PK, bool FP>
__device__ __forceinline__ void expand_tile(const uint64_t* __restrict__ sel_bm,
                                            const uint32_t* __restrict__ counts,
                                            const uint32_t* __restrict__ group_sums,
                                            const uint32_t* __restrict__ in_sel,
                                            uint32_t* __restrict__ out,
                                            uint32_t* __restrict__ out_count,
                                            const ExpandProj& pj, uint64_t* fp_groups,
                                            uint32_t tile, bool last) {
  __shared__ uint32_t s_ids[FTILE];
  __shared__ uint32_t s_wave[FWAVES];
  __shared__ uint32_t s_red[FWAVES];
  const int tid = threadIdx.x, wave = tid / WAVE, lane = tid % WAVE;
  const uint64_t t0 = (uint64_t)tile * FTILE;
  const uint32_t m = (uint32_t)(sel_bm[(t0 + (uint64_t)tid * 16) / 64] >> ((tid * 16) % 64)) & 0xFFFFu;
  const uint32_t base = group_sums ? tile_offset(counts, group_sums, tile, s_red) : counts[tile];
  const uint32_t c = (uint32_t)__popc(m);
  const uint32_t incl = wave_inclusive_u32(c);
  if (lane == WAVE - 1) s_wave[wave] = incl;
  __syncthreads();
  uint32_t pos = incl - c, total = 0;
#pragma unroll
  for (int w = 0; w < FWAVES; ++w) {
    if (w < wave) pos += s_wave[w];
    total += s_wave[w];
  }
  uint32_t mm = m;
  while (mm) {
    const int i = __ffs(mm) - 1;
    mm &= mm - 1;
    s_ids[pos++] = (uint32_t)tid * 16 + (uint32_t)i;  // row within the tile
  }
  __syncthreads();
  bool ovf = false;
  uint64_t h = 0;
  uint32_t i = tid;
  for (; i + 3 * FBLOCK < total; i += 4 * FBLOCK) {
    uint32_t r[4];
    uint64_t x[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint64_t k = t0 + s_ids[i + u * FBLOCK];
      r[u] = in_sel ? in_sel[k] : (uint32_t)k;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) x[u] = proj_load<PK>(pj, r[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint64_t o = (uint64_t)base + i + u * FBLOCK;
      out[o] = r[u];
      const uint64_t vb = proj_store<PK>(pj, o, x[u], ovf);
      if constexpr (FP) h += fp_term(o, r[u], PK ? vb : r[u]);
    }
  }
  for (; i < total; i += FBLOCK) {
    const uint64_t k = t0 + s_ids[i];
    const uint32_t r = in_sel ? in_sel[k] : (uint32_t)k;
    const uint64_t x = proj_load<PK>(pj, r);
    const uint64_t o = (uint64_t)base + i;
    out[o] = r;
    const uint64_t vb = proj_store<PK>(pj, o, x, ovf);
    if constexpr (FP) h += fp_term(o, r, PK ? vb : r);
  }
  if constexpr (PK == 1 || PK == 2) {
    if (__ballot(ovf) && lane == 0 && pj.ovf) atomicOr(pj.ovf, 1u);
  }
  if constexpr (FP) {
    __shared__ uint64_t s_fp[FWAVES];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) h += __shfl_xor(h, off, WAVE);
    if (lane == 0) s_fp[wave] = h;
    __syncthreads();
    if (tid == 0) {
      uint64_t t = 0;
#pragma unroll
      for (int w = 0; w < FWAVES; ++w) t += s_fp[w];
      if (t) atomicAdd((unsigned long long*)&fp_groups[(tile >> kGroupShift) * kFpStride],
                       (unsigned long long)t);
    }
  }
  if (group_sums && last && tid == 0) *out_count = base + total;
}

template <int PK, bool FP>
__global__ __launch_bounds__(FBLOCK) void k_expand(const uint64_t* __restrict__ sel_bm,
                                                   const uint32_t* __restrict__ counts,
                                                   const uint32_t* __restrict__ group_sums,
                                                   const uint32_t* __restrict__ in_sel,
                                                   uint32_t* __restrict__ out,
                                                   uint32_t* __restrict__ out_count,
                                                   ExpandProj pj, uint64_t* fp_groups) {
  expand_tile<PK, FP>(sel_bm, counts, group_sums, in_sel, out, out_count, pj, fp_groups,
                      blockIdx.x, blockIdx.x == gridDim.x - 1);
}

// ---------------------------------------------------------------------------
// Dense filter pass (no input selection, no probe or mark), the first of the
// two passes of a dense selection: the predicates are evaluated with loads
// of each column's own width (no 64-bit widening, so fewer registers and a
// higher occupancy than the general pass), the pass bits go to sel_bm, the
// tile's count to counts[tile] and to its group's sum.
// ---------------------------------------------------------------------------
struct DenseArgs {
  PredC preds[MTL_MAX_PREDS];
  uint32_t n_preds;
  uint32_t vec_ok;   // bit p: column p is aligned for vector loads
  uint32_t or_next;  // bit p: predicate p is OR-ed with predicate p + 1
  uint64_t n_rows;
  uint64_t* sel_bm;
  uint32_t* counts;
  uint32_t* group_sums;
};

// Clear the bits of the lane's 16 dense rows that fail p; T is the raw
// storage type, KIND the predicate's encoding (fixed at compile time so
// only this type's loads and compares are emitted).
template <typename T, uint32_t KIND, bool FULL>
__device__ __forceinline__ void pred_dense_t(const PredC& p, uint64_t k0, uint64_t n, bool vec,
                                             uint32_t& bits) {
  T x[FJ][4];
#pragma unroll
  for (int j = 0; j < FJ; ++j)
    load4<T, FULL>((const T*)p.col, k0 + (uint64_t)j * 256, n, vec, x[j]);
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (!pred_pass(enc_raw((uint64_t)x[j][v], KIND), p)) bits &= ~(1u << (4 * j + v));
}

template <bool FULL>
__device__ __forceinline__ void pred_dense_typed(const PredC& p, uint64_t k0, uint64_t n,
                                                 bool vec, uint32_t& bits) {
  switch (p.kind) {
    case PK_I32: pred_dense_t<uint32_t, PK_I32, FULL>(p, k0, n, vec, bits); break;
    case PK_F32: pred_dense_t<uint32_t, PK_F32, FULL>(p, k0, n, vec, bits); break;
    case PK_U8: pred_dense_t<uint8_t, PK_U8, FULL>(p, k0, n, vec, bits); break;
    case PK_I64: pred_dense_t<uint64_t, PK_I64, FULL>(p, k0, n, vec, bits); break;
    default: pred_dense_t<uint64_t, PK_F64, FULL>(p, k0, n, vec, bits); break;
  }
}

__global__ __launch_bounds__(FBLOCK) void k_filter_dense(DenseArgs a) {
  __shared__ uint32_t s_wave[FWAVES];
  const int tid = threadIdx.x, wave = tid / WAVE, lane = tid % WAVE;
  const uint32_t tile = blockIdx.x;
  const uint64_t n = a.n_rows;
  const uint64_t seg = (uint64_t)tile * FTILE + (uint64_t)wave * (FJ * 256);
  const uint64_t k0 = seg + (uint64_t)lane * 4;
  uint32_t bits = 0;
#pragma unroll
  for (int j = 0; j < FJ; ++j)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (k0 + (uint64_t)j * 256 + v < n) bits |= 1u << (4 * j + v);
  // a tile inside the table with every column aligned: straight-line
  // streaming loads (block-uniform choice)
  const bool full = (uint64_t)(tile + 1) * FTILE <= n &&
                    a.vec_ok == (a.n_preds >= 32 ? ~0u : (1u << a.n_preds) - 1u);
  if (!a.or_next) {
    for (uint32_t p = 0; p < a.n_preds; ++p) {
      if (full) pred_dense_typed<true>(a.preds[p], k0, n, true, bits);
      else pred_dense_typed<false>(a.preds[p], k0, n, (a.vec_ok >> p) & 1u, bits);
    }
  } else {
    // a conjunction of disjunctions: each run of OR-ed predicates collects
    // its rows in `any`, and-ed into bits at the run's last predicate
    const uint32_t live = bits;
    uint32_t any = 0;
    for (uint32_t p = 0; p < a.n_preds; ++p) {
      uint32_t m = live;
      if (full) pred_dense_typed<true>(a.preds[p], k0, n, true, m);
      else pred_dense_typed<false>(a.preds[p], k0, n, (a.vec_ok >> p) & 1u, m);
      any |= m;
      if (!((a.or_next >> p) & 1u)) {
        bits &= any;
        any = 0;
      }
    }
  }
  store_pass_bits(a.sel_bm, seg, bits, lane);
  const uint32_t c = wave_sum_u32((uint32_t)__popc(bits));
  if (lane == 0) s_wave[wave] = c;
  __syncthreads();
  if (tid == 0) {
    uint32_t t = 0;
#pragma unroll
    for (int w = 0; w < FWAVES; ++w) t += s_wave[w];
    a.counts[tile] = t;
    if (t) atomicAdd(&a.group_sums[(tile >> kGroupShift) * kGroupStride], t);
  }
}

void scan_counts(uint32_t* counts, uint32_t tiles, uint32_t* total, hipStream_t st) {
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(SBLOCK), 0, st, counts, tiles, total);
}

// Kernel variant of a projection: 1 int32, 2 int64, 3 float, 4 double.
// DATE and TIMESTAMP take only + and - (of days, of microseconds); DuckDB
// has no DATE * INTEGER.
static int32_t proj_kind(int32_t type, int32_t op, int* pk) {
  if (op < 0 || op > 3) return set_error(MTL_ERR_INVALID_ARGUMENT, "project op %d", op);
  switch (type) {
    case MTL_INTEGER: *pk = 1; return MTL_OK;
    case MTL_BIGINT:
    case MTL_DECIMAL: *pk = 2; return MTL_OK;
    case MTL_DATE:
    case MTL_TIMESTAMP:
      if (op == MTL_PROJ_MUL || op == MTL_PROJ_SUBFROM)
        return set_error(MTL_UNSUPPORTED, "%s supports only + and -",
                         type == MTL_DATE ? "DATE" : "TIMESTAMP");
      *pk = type == MTL_DATE ? 1 : 2;
      return MTL_OK;
    case MTL_FLOAT: *pk = 3; return MTL_OK;
    case MTL_DOUBLE: *pk = 4; return MTL_OK;
    default: return set_error(MTL_UNSUPPORTED, "projection on type %d", type);
  }
}

int32_t launch_compact(const uint64_t* sel_bm, uint32_t* tile_counts, uint32_t tiles,
                       const uint32_t* in_sel, uint32_t* out_sel, uint32_t* out_count,
                       hipStream_t st, const CompactProj* proj, const uint32_t* group_sums,
                       uint64_t* fp_groups) {
  if (!group_sums) {
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(SBLOCK), 0, st, tile_counts, tiles,
                       out_count);
    MTL_HIP(hipGetLastError());
  }
  ExpandProj pj{};
  int pk = 0;
  if (proj) {
    pj.col = proj->col;
    pj.op = proj->op;
    pj.scalar = proj->scalar;
    pj.out = proj->out;
    pj.ovf = proj->ovf;
    MTL_TRY(proj_kind(proj->type, proj->op, &pk));
  }
#define MTL_EXPAND(K, F)                                                                        \
  hipLaunchKernelGGL((k_expand<K, F>), dim3(tiles), dim3(FBLOCK), 0, st, sel_bm, tile_counts,  \
                     group_sums, in_sel, out_sel, out_count, pj, fp_groups)
  if (fp_groups) {
    switch (pk) {
      case 1: MTL_EXPAND(1, true); break;
      case 2: MTL_EXPAND(2, true); break;
      case 3: MTL_EXPAND(3, true); break;
      case 4: MTL_EXPAND(4, true); break;
      default: MTL_EXPAND(0, true); break;
    }
  } else {
    switch (pk) {
      case 1: MTL_EXPAND(1, false); break;
      case 2: MTL_EXPAND(2, false); break;
      case 3: MTL_EXPAND(3, false); break;
      case 4: MTL_EXPAND(4, false); break;
      default: MTL_EXPAND(0, false); break;
    }
  }
#undef MTL_EXPAND
  MTL_HIP(hipGetLastError());
  return MTL_OK;
}

static int32_t launch_filter_dense(DenseArgs a, uint32_t tiles, hipStream_t st) {
  a.vec_ok = 0;
  for (uint32_t p = 0; p < a.n_preds; ++p) {
    const uint32_t w = pred_width(a.preds[p].kind);
    if ((uintptr_t)a.preds[p].col % (w == 1 ? 4 : 16) == 0) a.vec_ok |= 1u << p;
  }
  hipLaunchKernelGGL(k_filter_dense, dim3(tiles), dim3(FBLOCK), 0, st, a);
  MTL_HIP(hipGetLastError());
  return MTL_OK;
}

uint32_t filter_tiles(uint64_t n) { return (uint32_t)ceil_div(n ? n : 1, FTILE); }

int32_t launch_filter(FilterArgs a, uint64_t capacity, hipStream_t st) {
  uint32_t tiles = filter_tiles(capacity);
  a.vec_ok = 0;
  for (uint32_t p = 0; p < a.n_preds; ++p) {
    int w = (int)dev::pred_width(a.preds[p].kind);
    uintptr_t addr = (uintptr_t)a.preds[p].col;
    int need = w == 1 ? 4 : 16;
    if (addr % need == 0) a.vec_ok |= 1u << p;
  }
  if (a.in_sel && ((uintptr_t)a.in_sel % 16) != 0)
    return set_error(MTL_ERR_INVALID_ARGUMENT,
                     "selection vectors must be 16-byte aligned");
  auto key_vec = [](const void* col, int32_t type) -> uint32_t {
    if (!col) return 0;
    int w = type_width(type);
    return ((uintptr_t)col % (w == 1 ? 4 : 16)) == 0 ? 1u : 0u;
  };
  a.probe_vec = key_vec(a.probe_key, a.probe_key_type);
  a.mark_vec = key_vec(a.mark_key, a.mark_key_type);
  auto narrow_key = [](const void* col, int32_t type, uint64_t span) {
    return !col || (type_width(type) <= 4 && span < (1ull << 32));
  };
  const bool narrow = !a.in_sel && a.n_preds <= 1 &&
                      (a.n_preds == 0 || dev::pred_width(a.preds[0].kind) <= 4) &&
                      (a.probe_bm || a.mark_bm) &&
                      narrow_key(a.probe_bm ? a.probe_key : nullptr, a.probe_key_type,
                                 a.probe_span) &&
                      narrow_key(a.mark_bm ? a.mark_key : nullptr, a.mark_key_type, a.mark_span);
  // probe-bound passes (a large table, a probe bitmap of 64K keys or more,
  // pass bits and marks only) take the persistent kernel with the bitmap's
  // low words in LDS
  const int pk = a.n_preds ? (int)a.preds[0].kind : -1;
  const bool lds_probe = narrow && a.probe_bm && a.mark_bm && a.sel_bm && !a.tile_counts &&
                         !a.group_sums && a.n_rows >= (1u << 20) &&
                         a.probe_span >= (1u << 16) && ((uintptr_t)a.probe_bm % 16) == 0 &&
                         type_width(a.probe_key_type) == 4 &&
                         type_width(a.mark_key_type) == 4 &&
                         (pk < 0 || pk == (int)PK_I32 || pk == (int)PK_F32);
  if (lds_probe) {
    const uint64_t words32 = ceil_div(a.probe_span, 64) * 2;
    const uint32_t lds_words = (uint32_t)std::min<uint64_t>(kLdsProbeWords, words32 & ~3ull);
    const uint64_t segs = (uint64_t)tiles * (FTILE / (FJ * 256));
    const uint32_t grid = (uint32_t)std::min<uint64_t>(cu_count(), ceil_div(segs, PBLOCK / WAVE));
    auto fits32 = [](int64_t min, uint64_t span) {
      return min >= 0 && (uint64_t)min + span <= (1ull << 31);
    };
    const bool k32 = fits32(a.probe_min, a.probe_span) && fits32(a.mark_min, a.mark_span);
#define MTL_PROBE(PK, K32) \
  hipLaunchKernelGGL((k_filter_probe<PK, K32>), dim3(grid), dim3(PBLOCK), 0, st, a, lds_words)
#define MTL_PROBE_K(PK) \
  if (k32) MTL_PROBE(PK, true); else MTL_PROBE(PK, false)
    if (pk == (int)PK_I32) { MTL_PROBE_K((int)PK_I32); }
    else if (pk == (int)PK_F32) { MTL_PROBE_K((int)PK_F32); }
    else { MTL_PROBE_K(-1); }
#undef MTL_PROBE_K
#undef MTL_PROBE
  } else if (a.in_sel) {
    hipLaunchKernelGGL((k_filter<true, false>), dim3(tiles), dim3(FBLOCK), 0, st, a);
  } else if (narrow)
    hipLaunchKernelGGL((k_filter<false, true>), dim3(tiles), dim3(FBLOCK), 0, st, a);
  else hipLaunchKernelGGL((k_filter<false, false>), dim3(tiles), dim3(FBLOCK), 0, st, a);
  MTL_HIP(hipGetLastError());
  return MTL_OK;
}

// ---------------------------------------------------------------------------
// Mask output (compare_*_values counterpart).
// ---------------------------------------------------------------------------
struct CompareArgs {
  PredC preds[MTL_MAX_PREDS];
  uint32_t n_preds;
  uint64_t n;
  uint8_t* out;
};

__global__ void k_compare(CompareArgs a) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    bool ok = true;
    for (uint32_t p = 0; p < a.n_preds && ok; ++p) {
      const PredC& c = a.preds[p];
      ok = pred_pass(enc_raw(load_raw(c.col, pred_width(c.kind), i), c.kind), c);
    }
    a.out[i] = ok ? 1 : 0;
  }
}

// ---------------------------------------------------------------------------
// Gather through a selection (count on the device).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_gather(const T* __restrict__ col, const uint32_t* __restrict__ sel,
                         const uint32_t* count, uint64_t cap, T* __restrict__ out) {
  const uint64_t n = dev_count(count, cap);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = col[sel ? sel[i] : i];
}

// ---------------------------------------------------------------------------
// Scalar projection with DuckDB overflow semantics.
// ---------------------------------------------------------------------------
// Each thread handles PU elements per grid stride, with all their selection
// loads, then all their column gathers, issued before the first use.
constexpr int PU = 4;
template <typename T, bool IS_FP>
__global__ void k_project(const T* __restrict__ col, const uint32_t* __restrict__ sel,
                          const uint32_t* count, uint64_t cap, int op, T s,
                          T* __restrict__ out, uint32_t* ovf_flag) {
  const uint64_t n = dev_count(count, cap);
  const uint64_t G = (uint64_t)gridDim.x * blockDim.x;
  bool ovf = false;
  for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += G * PU) {
    uint32_t r[PU];
    T x[PU];
#pragma unroll
    for (int k = 0; k < PU; ++k) {
      const uint64_t i = i0 + k * G;
      r[k] = i < n ? (sel ? sel[i] : (uint32_t)i) : 0u;
    }
#pragma unroll
    for (int k = 0; k < PU; ++k)
      if (i0 + k * G < n) x[k] = col[r[k]];
#pragma unroll
    for (int k = 0; k < PU; ++k) {
      const uint64_t i = i0 + k * G;
      if (i < n) {
        if constexpr (IS_FP) out[i] = proj_fp<T>(x[k], s, op);
        else out[i] = proj_int<T>(x[k], s, op, ovf);
      }
    }
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

static int32_t filter_check(const mtl_pred* preds, uint32_t n_preds, uint64_t n_rows,
                            const uint32_t* in_sel, uint64_t in_capacity, uint32_t* out_sel,
                            uint32_t* out_count, uint64_t* cap) {
  if (n_preds > MTL_MAX_PREDS)
    return set_error(MTL_UNSUPPORTED, "more than %d predicates", MTL_MAX_PREDS);
  if (!out_sel || !out_count)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "out_sel/out_count are NULL");
  if (n_preds && !preds) return set_error(MTL_ERR_INVALID_ARGUMENT, "preds is NULL");
  for (uint32_t p = 0; p < n_preds; ++p)
    if (!type_width(preds[p].type) || preds[p].type == MTL_HUGEINT ||
        preds[p].op < 0 || preds[p].op > 4 || !preds[p].column)
      return set_error(MTL_UNSUPPORTED, "predicate %u: unsupported type/op", p);
  *cap = in_sel ? in_capacity : n_rows;
  if (*cap > 0xFFFFFFFFull)
    return set_error(MTL_UNSUPPORTED, "more than 2^32-1 rows");
  return MTL_OK;
}

// Device scratch of one filter over `tiles` tiles: pass bits, tile counts,
// a clearing word block, group sums (and fingerprint sums).
struct FilterScratch {
  void *bm, *counts, *ctr, *gsum;
};

// The filter pass and the expansion on s with the given scratch.  prof:
// record kernel timings (creates events when profiling is on).
static int32_t filter_enqueue(const mtl_pred* preds, uint32_t n_preds, uint64_t cap,
                              const uint32_t* in_sel, const uint32_t* in_count,
                              uint32_t* out_sel, uint32_t* out_count, const CompactProj* proj,
                              ScanExtras* ex, const FilterScratch& sc, uint64_t* fpg,
                              uint32_t tiles, uint32_t groups, bool prof, mtl_stream* s,
                              uint32_t or_next = 0) {
  if (or_next && in_sel)
    return set_error(MTL_UNSUPPORTED, "OR-ed predicates over an input selection");
  ZeroList zl = ex ? ex->zero : ZeroList{};
  if (zl.n > 5) return set_error(MTL_ERR_INVALID_ARGUMENT, "too many regions to clear");
  zl.add((uint32_t*)sc.ctr, 64);
  zl.add((uint32_t*)sc.gsum, (uint64_t)groups * kGroupStride);
  if (fpg) zl.add((uint32_t*)fpg, (uint64_t)groups * kFpStride * 2);
  MTL_TRY(launch_zero(zl, s->stream));
  if (!in_sel) {
    DenseArgs a{};
    for (uint32_t p = 0; p < n_preds; ++p) a.preds[p] = compile_pred(preds[p]);
    a.n_preds = n_preds;
    a.or_next = n_preds ? or_next & ((1u << (n_preds - 1)) - 1u) : 0u;  // last ends a run
    a.n_rows = cap;
    a.sel_bm = (uint64_t*)sc.bm;
    a.counts = (uint32_t*)sc.counts;
    a.group_sums = (uint32_t*)sc.gsum;
    ProfScope ps(s, "k_filter", prof);
    MTL_TRY(launch_filter_dense(a, tiles, s->stream));
  } else {
    FilterArgs a{};
    for (uint32_t p = 0; p < n_preds; ++p) a.preds[p] = compile_pred(preds[p]);
    a.n_preds = n_preds;
    a.n_rows = cap;
    a.in_sel = in_sel;
    a.in_count = in_count;
    a.sel_bm = (uint64_t*)sc.bm;
    a.tile_counts = (uint32_t*)sc.counts;
    a.group_sums = (uint32_t*)sc.gsum;
    a.err = (uint32_t*)sc.ctr;
    ProfScope ps(s, "k_filter", prof);
    MTL_TRY(launch_filter(a, cap, s->stream));
  }
  ProfScope ps(s, "k_compact", prof);
  return launch_compact((const uint64_t*)sc.bm, (uint32_t*)sc.counts, tiles, in_sel, out_sel,
                        out_count, s->stream, proj, (const uint32_t*)sc.gsum, fpg);
}

static int32_t filter_impl(const mtl_pred* preds, uint32_t n_preds, uint64_t n_rows,
                           const uint32_t* in_sel, const uint32_t* in_count,
                           uint64_t in_capacity, uint32_t* out_sel, uint32_t* out_count,
                           const CompactProj* proj, ScanExtras* ex, mtl_stream* s,
                           uint32_t or_next = 0) {
  uint64_t cap;
  MTL_TRY(filter_check(preds, n_preds, n_rows, in_sel, in_capacity, out_sel, out_count, &cap));
  MTL_TRY(resolve_stream(s, &s));
  std::lock_guard<std::mutex> g(s->mu);
  s->arena.reset();
  const uint32_t tiles = filter_tiles(cap);
  const uint32_t groups = (uint32_t)ceil_div(tiles, 1u << kGroupShift);
  FilterScratch sc;
  MTL_TRY(s->arena.alloc((size_t)tiles * FTILE / 8, &sc.bm));
  MTL_TRY(s->arena.alloc((size_t)tiles * 4, &sc.counts));
  MTL_TRY(s->arena.alloc(256, &sc.ctr));
  MTL_TRY(s->arena.alloc((size_t)groups * kGroupStride * 4, &sc.gsum));
  uint64_t* fpg = nullptr;
  if (ex && ex->fingerprint) {
    void* p;
    MTL_TRY(s->arena.alloc((size_t)groups * kFpStride * 8, &p));
    fpg = (uint64_t*)p;
    ex->fp_groups = fpg;
    ex->fp_n = groups;
  }
  return filter_enqueue(preds, n_preds, cap, in_sel, in_count, out_sel, out_count, proj, ex,
                        sc, fpg, tiles, groups, true, s, or_next);
}

}  // extern "C"

namespace mtl {

// ---------------------------------------------------------------------------
// The filter without allocation (mtl_filter_noalloc): its pass bits, tile
// counts, clearing words and group sums live in this file-scope device
// array, one per GPU the library runs on, held until the process ends.
// ---------------------------------------------------------------------------
constexpr uint64_t kNoAllocScratchBytes = 16ull << 20;
constexpr uint64_t kNoAllocMaxRows = 100000000ull;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// The device address of the current GPU's `scratch`, looked up once per
// device (in host code the array's name is not a device address).
static int32_t noalloc_scratch(void** out) {
  static std::mutex mu;
  static void* ptr[64] = {};
  int dev = 0;
  MTL_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return set_error(MTL_ERR_DEVICE_UNAVAILABLE, "device %d", dev);
  std::lock_guard<std::mutex> g(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(MTL_ERR_PROCESSING, "processing error: hipGetSymbolAddress(scratch): %s",
                       hipGetErrorString(e));
    ptr[dev] = p;
  }
  *out = ptr[dev];
  return MTL_OK;
}

}  // namespace mtl

extern "C" {

int32_t mtl_filter_noalloc(const mtl_pred* preds, uint32_t n_preds, uint64_t n_rows,
                           const uint32_t* in_sel, const uint32_t* in_count,
                           uint64_t in_capacity, uint32_t* out_sel, uint32_t* out_count,
                           mtl_stream* s) {
  uint64_t cap;
  MTL_TRY(filter_check(preds, n_preds, n_rows, in_sel, in_capacity, out_sel, out_count, &cap));
  if (cap > kNoAllocMaxRows)
    return set_error(MTL_ERR_INVALID_ARGUMENT,
                     "mtl_filter_noalloc takes at most %llu rows (%llu given): use mtl_filter",
                     (unsigned long long)kNoAllocMaxRows, (unsigned long long)cap);
  if (!s) {
    s = runtime().default_stream;  // creating it would allocate
    if (!s)
      return set_error(MTL_ERR_NOT_INITIALIZED,
                       "mtl_filter_noalloc: pass a stream, or call mtl_init first");
  }
  void* base;
  MTL_TRY(noalloc_scratch(&base));
  const uint32_t tiles = filter_tiles(cap);
  const uint32_t groups = (uint32_t)ceil_div(tiles, 1u << kGroupShift);
  auto up = [](uint64_t b) { return (b + 255) & ~uint64_t(255); };
  FilterScratch sc;
  char* at = (char*)base;
  sc.bm = at;
  at += up((uint64_t)tiles * FTILE / 8);
  sc.counts = at;
  at += up((uint64_t)tiles * 4);
  sc.ctr = at;
  at += 256;
  sc.gsum = at;
  at += up((uint64_t)groups * kGroupStride * 4);
  if ((uint64_t)(at - (char*)base) > kNoAllocScratchBytes)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "mtl_filter_noalloc: %llu rows exceed the scratch",
                     (unsigned long long)cap);
  std::lock_guard<std::mutex> g(s->mu);
  return filter_enqueue(preds, n_preds, cap, in_sel, in_count, out_sel, out_count, nullptr,
                        nullptr, sc, nullptr, tiles, groups, false, s);
}

int32_t mtl_filter(const mtl_pred* preds, uint32_t n_preds, uint64_t n_rows,
                   const uint32_t* in_sel, const uint32_t* in_count,
                   uint64_t in_capacity, uint32_t* out_sel, uint32_t* out_count,
                   mtl_stream* s) {
  return filter_impl(preds, n_preds, n_rows, in_sel, in_count, in_capacity, out_sel,
                     out_count, nullptr, nullptr, s);
}

int32_t mtl_filter_or(const mtl_pred* preds, uint32_t n_preds, uint32_t or_next,
                      uint64_t n_rows, uint32_t* out_sel, uint32_t* out_count, mtl_stream* s) {
  return filter_impl(preds, n_preds, n_rows, nullptr, nullptr, 0, out_sel, out_count, nullptr,
                     nullptr, s, or_next);
}

int32_t mtl_filter_project(const mtl_pred* preds, uint32_t n_preds, uint64_t n_rows,
                           const uint32_t* in_sel, const uint32_t* in_count,
                           uint64_t in_capacity, const void* proj_column, int32_t proj_type,
                           int32_t op, int64_t scalar, uint32_t* out_sel,
                           uint32_t* out_count, void* out_proj, uint32_t* overflow_flag,
                           mtl_stream* s) {
  int pk = 0;
  MTL_TRY(proj_kind(proj_type, op, &pk));
  if (!proj_column || !out_proj)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "proj_column/out_proj are NULL");
  CompactProj pj{proj_column, proj_type, op, scalar, out_proj, overflow_flag};
  return filter_impl(preds, n_preds, n_rows, in_sel, in_count, in_capacity, out_sel,
                     out_count, &pj, nullptr, s);
}

int32_t mtl_compare(const mtl_pred* preds, uint32_t n_preds, uint64_t n_rows,
                    uint8_t* out_mask, mtl_stream* s) {
  if (n_preds > MTL_MAX_PREDS)
    return set_error(MTL_UNSUPPORTED, "more than %d predicates", MTL_MAX_PREDS);
  MTL_TRY(resolve_stream(s, &s));
  CompareArgs a{};
  for (uint32_t p = 0; p < n_preds; ++p) a.preds[p] = compile_pred(preds[p]);
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
  int pk = 0;
  MTL_TRY(proj_kind(type, op, &pk));
  MTL_TRY(resolve_stream(s, &s));
  if (capacity == 0) return MTL_OK;
  if (!column || !out) return set_error(MTL_ERR_INVALID_ARGUMENT, "NULL argument");
  dim3 g(grid_for(capacity)), b(256);
  switch (pk) {
    case 1:
      hipLaunchKernelGGL((k_project<int32_t, false>), g, b, 0, s->stream,
                         (const int32_t*)column, sel, count, capacity, op,
                         (int32_t)scalar, (int32_t*)out, overflow_flag);
      break;
    case 2:
      hipLaunchKernelGGL((k_project<int64_t, false>), g, b, 0, s->stream,
                         (const int64_t*)column, sel, count, capacity, op, scalar,
                         (int64_t*)out, overflow_flag);
      break;
    case 3:
      hipLaunchKernelGGL((k_project<float, true>), g, b, 0, s->stream,
                         (const float*)column, sel, count, capacity, op,
                         (float)__builtin_bit_cast(double, scalar), (float*)out,
                         overflow_flag);
      break;
    default:
      hipLaunchKernelGGL((k_project<double, true>), g, b, 0, s->stream,
                         (const double*)column, sel, count, capacity, op,
                         __builtin_bit_cast(double, scalar), (double*)out,
                         overflow_flag);
      break;
  }
  MTL_HIP(hipGetLastError());
  return MTL_OK;
}

}  // extern "C"

namespace mtl {

int32_t filter_project_ex(const mtl_pred* preds, uint32_t n_preds, uint64_t n_rows,
                          const CompactProj* proj, uint32_t* out_sel, uint32_t* out_count,
                          ScanExtras* ex, mtl_stream* s) {
  if (proj) {
    int pk = 0;
    MTL_TRY(proj_kind(proj->type, proj->op, &pk));
    if (!proj->col || !proj->out)
      return set_error(MTL_ERR_INVALID_ARGUMENT, "proj_column/out_proj are NULL");
  }
  return filter_impl(preds, n_preds, n_rows, nullptr, nullptr, 0, out_sel, out_count, proj, ex,
                     s);
}

}  // namespace mtl

// ---------------------------------------------------------------------------
// Shared scan of a batch of queries.  k_mq_filter loads each tile of the
// shared predicate column once and, query by query, writes the pass bits,
// the tile count and the group sum (the same layout as the single-query
// filter pass); k_mq_expand stages the tile's projection column in LDS once
// (when enough of its rows pass, summed over the queries) and expands every
// query's selection, projection and fingerprint from it, with each query's
// output offset computed from its own counts and group sums.
// ---------------------------------------------------------------------------
namespace mtl {

struct MQScanQ {
  PredC pred;
  uint64_t n;
  uint32_t tiles;
  int32_t op;
  int64_t scalar;
  uint32_t* out_sel;
  void* out_proj;
  uint64_t* bm;
  uint32_t* counts;  // per tile; then, scanned, each tile's output offset
  uint64_t* fp;
  uint32_t* words;  // [0] count, [1] overflow, [2..3] fingerprint
};

struct MQScanArgs {
  const void* col;   // shared predicate column
  const void* pcol;  // shared projection column (nullptr: none)
  uint64_t n_max;    // rows of the largest query
  uint32_t nq;
  const MQScanQ* q;
};

// Shared passes: one block per tile, the tile's column staged in LDS once;
// each wave takes one query at a time, so the queries of a tile run in
// parallel on the block's waves with no block barrier between them.
// Within a wave, lane l covers rows 64 i + l of strip i (i = 0..63), so an
// LDS read of a strip is conflict-free and a ballot over the strip is the
// strip's 64-bit pass-bit word, stored by lane i.
constexpr int MQ_BLOCK = 256;
constexpr int MQ_WAVES = MQ_BLOCK / WAVE;

// The tile's values of a column of T into LDS (rows past n read as 0): a
// full tile of an aligned column with 16-byte loads, all in flight together.
template <typename T>
__device__ __forceinline__ void mq_stage(T* s_val, const T* col, uint64_t t0, uint64_t n) {
  constexpr uint32_t PER = 16 / sizeof(T);  // values per 16-byte load
  constexpr uint32_t NV = FTILE / PER / MQ_BLOCK;  // loads per thread
  if (t0 + FTILE <= n && ((uintptr_t)col % 16) == 0) {
    uint4 v[NV];
#pragma unroll
    for (uint32_t u = 0; u < NV; ++u)
      v[u] = ld_stream((const uint4*)(col + t0) + threadIdx.x + u * MQ_BLOCK);
#pragma unroll
    for (uint32_t u = 0; u < NV; ++u) ((uint4*)s_val)[threadIdx.x + u * MQ_BLOCK] = v[u];
    return;
  }
  for (uint32_t i = threadIdx.x; i < FTILE; i += MQ_BLOCK) {
    const uint64_t r = t0 + i;
    s_val[i] = r < n ? col[r] : T(0);
  }
}

template <typename T, uint32_t KIND>
__global__ __launch_bounds__(MQ_BLOCK) void k_mq_filter(MQScanArgs a) {
  __shared__ __align__(16) T s_val[FTILE];
  const uint32_t tile = blockIdx.x;
  const uint64_t t0 = (uint64_t)tile * FTILE;
  mq_stage<T>(s_val, (const T*)a.col, t0, a.n_max);
  __syncthreads();
  const int wave = threadIdx.x / WAVE, lane = lane_id();
  for (uint32_t q = wave; q < a.nq; q += MQ_WAVES) {
    const MQScanQ& Q = a.q[q];
    if (tile >= Q.tiles) continue;  // wave-uniform
    const uint64_t n = Q.n;
    const PredC p = Q.pred;  // in registers for the strip loop
    // rows of this lane below n: strips i < full are whole, strip `full`
    // is partial
    const uint64_t rows = n - t0 < (uint64_t)FTILE ? n - t0 : (uint64_t)FTILE;
    uint64_t mine = 0;
    uint32_t cnt = 0;
#pragma unroll 1
    for (int i0 = 0; i0 < 64; i0 += 16) {
      // sixteen strips' values read before any is compared: no branch
      // between the LDS reads, so they are all in flight together
      T v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) v[u] = s_val[(uint32_t)(i0 + u) * 64 + lane];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const uint32_t li = (uint32_t)(i0 + u) * 64 + lane;
        bool ok;
        if constexpr (KIND == PK_I32) ok = pred_pass_i32((uint32_t)v[u], p);
        else ok = pred_pass(enc_raw((uint64_t)v[u], KIND), p);
        const uint64_t w = __ballot((li < rows) & ok);
        cnt += (uint32_t)__popcll(w);
        mine = lane == i0 + u ? w : mine;
      }
    }
    Q.bm[(uint64_t)tile * 64 + lane] = mine;
    if (lane == 0) Q.counts[tile] = cnt;
  }
}

// Block q scans query q's tile counts in place into exclusive output
// offsets and writes the total (the query's row count) to words[0].
__global__ __launch_bounds__(1024) void k_mq_offsets(const MQScanQ* qs) {
  __shared__ uint32_t s_wave[1024 / WAVE];
  const MQScanQ& Q = qs[blockIdx.x];
  const uint32_t T = Q.tiles;
  const uint32_t per = (T + 1023) / 1024;
  const uint32_t b0 = threadIdx.x * per, b1 = min(T, b0 + per);
  uint32_t sum = 0;
  for (uint32_t i = b0; i < b1; ++i) sum += Q.counts[i];
  const uint32_t incl = wave_inclusive_u32(sum);
  const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  if (lane == WAVE - 1) s_wave[wave] = incl;
  __syncthreads();
  uint32_t run = incl - sum, total = 0;
  for (int w = 0; w < 1024 / WAVE; ++w) {
    if (w < wave) run += s_wave[w];
    total += s_wave[w];
  }
  for (uint32_t i = b0; i < b1; ++i) {
    const uint32_t c = Q.counts[i];
    Q.counts[i] = run;
    run += c;
  }
  if (threadIdx.x == 0) Q.words[0] = total;
}

template <int PK>
__global__ __launch_bounds__(MQ_BLOCK) void k_mq_expand(MQScanArgs a) {
  using VT = typename std::conditional<PK == 2 || PK == 4, uint64_t, uint32_t>::type;
  __shared__ __align__(16) VT s_val[PK ? FTILE : 2];
  const uint32_t tile = blockIdx.x;
  const uint64_t t0 = (uint64_t)tile * FTILE;
  if constexpr (PK != 0) {
    mq_stage<VT>(s_val, (const VT*)a.pcol, t0, a.n_max);
    __syncthreads();
  }
  const int wave = threadIdx.x / WAVE, lane = lane_id();
  const uint64_t lt = lanemask_lt();
  // each wave takes every MQ_WAVES-th query; the next query's pass word and
  // output offset are loaded before the current one is expanded
  uint32_t q = wave;
  uint64_t nx_words = 0;
  uint32_t nx_off = 0;
  if (q < a.nq && tile < a.q[q].tiles) {
    nx_words = a.q[q].bm[(uint64_t)tile * 64 + lane];
    nx_off = a.q[q].counts[tile];
  }
#pragma unroll 1
  for (; q < a.nq; q += MQ_WAVES) {
    const MQScanQ& Q = a.q[q];
    const uint64_t mine = nx_words;
    const uint32_t off = nx_off;
    const uint32_t qn = q + MQ_WAVES;
    if (qn < a.nq && tile < a.q[qn].tiles) {
      nx_words = a.q[qn].bm[(uint64_t)tile * 64 + lane];
      nx_off = a.q[qn].counts[tile];
    }
    if (tile >= Q.tiles) continue;  // wave-uniform
    const uint32_t c = (uint32_t)__popcll(mine);
    const uint32_t incl = wave_inclusive_u32(c);
    const uint32_t excl = incl - c;
    const ExpandProj pj{a.pcol, Q.op, Q.scalar, Q.out_proj, nullptr};
    uint32_t* const out_sel = Q.out_sel;
    bool ovf = false;
    uint64_t h = 0;
    const uint32_t mlo = (uint32_t)mine, mhi = (uint32_t)(mine >> 32);
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, WAVE - 1);
    if (total * 2 < FTILE) {
      // fewer than half the rows pass: one output position per lane, its
      // strip found by a binary search over the lanes' exclusive counts and
      // its row by selecting the k-th set bit of the strip's word
      for (uint32_t p0 = 0; p0 < total; p0 += WAVE) {
        const uint32_t p = p0 + (uint32_t)lane;
        uint32_t s = 0;
#pragma unroll
        for (uint32_t step = 32; step > 0; step >>= 1) {
          const uint32_t e = (uint32_t)__shfl((int)excl, (int)(s + step), WAVE);
          s = e <= p ? s + step : s;
        }
        uint32_t k = p - (uint32_t)__shfl((int)excl, (int)s, WAVE);
        const uint32_t wl = (uint32_t)__shfl((int)mlo, (int)s, WAVE);
        const uint32_t wh = (uint32_t)__shfl((int)mhi, (int)s, WAVE);
        const uint32_t cl = (uint32_t)__popc(wl);
        uint32_t x = k < cl ? wl : wh;
        uint32_t bit = k < cl ? 0u : 32u;
        k = k < cl ? k : k - cl;
#pragma unroll
        for (uint32_t wdt = 16; wdt > 0; wdt >>= 1) {
          const uint32_t cw = (uint32_t)__popc(x & ((1u << wdt) - 1u));
          const bool up = k >= cw;
          k = up ? k - cw : k;
          x = up ? x >> wdt : x;
          bit = up ? bit + wdt : bit;
        }
        const uint32_t li = s * 64 + bit;
        uint64_t xv = 0;
        if constexpr (PK != 0) xv = (uint64_t)s_val[li & (FTILE - 1)];
        if (p < total) {
          const uint32_t r = (uint32_t)(t0 + li);
          const uint64_t o = (uint64_t)off + p;
          out_sel[o] = r;
          const uint64_t vb = proj_store<PK>(pj, o, xv, ovf);
          h += fp_term(o, r, PK ? vb : r);
        }
      }
    } else {
      // half or more pass: strip by strip, every lane one row of the strip
      for (int i = 0; i < 64; ++i) {
        // strip i's word and output base from lane i (scalar reads)
        const uint64_t w = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)mhi, i) << 32) |
                           (uint32_t)__builtin_amdgcn_readlane((int)mlo, i);
        if (!w) continue;  // uniform
        const uint32_t base = (uint32_t)__builtin_amdgcn_readlane((int)excl, i);
        const uint32_t li = (uint32_t)i * 64 + lane;
        uint64_t x = 0;
        if constexpr (PK != 0) x = (uint64_t)s_val[li];  // before the branch: no wait inside it
        if ((w >> lane) & 1ull) {
          const uint32_t r = (uint32_t)(t0 + li);
          const uint64_t o = (uint64_t)off + base + (uint32_t)__popcll(w & lt);
          out_sel[o] = r;
          const uint64_t vb = proj_store<PK>(pj, o, x, ovf);
          h += fp_term(o, r, PK ? vb : r);
        }
      }
    }
    if constexpr (PK == 1 || PK == 2) {
      if (__ballot(ovf) && lane == 0) atomicOr(&Q.words[1], 1u);
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) h += __shfl_xor(h, d, WAVE);
    if (lane == 0 && h)
      atomicAdd((unsigned long long*)&Q.fp[(tile >> kGroupShift) * kFpStride],
                (unsigned long long)h);
  }
}

// Block q adds up query q's per-group fingerprint sums.
__global__ __launch_bounds__(1024) void k_mq_fp_sum(const MQScanQ* qs) {
  __shared__ uint64_t s_part[1024 / WAVE];
  const MQScanQ& Q = qs[blockIdx.x];
  const uint32_t groups = (Q.tiles + (1u << kGroupShift) - 1) >> kGroupShift;
  uint64_t t = 0;
  for (uint32_t i = threadIdx.x; i < groups; i += blockDim.x) t += Q.fp[(uint64_t)i * kFpStride];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, WAVE);
  if (threadIdx.x % WAVE == 0) s_part[threadIdx.x / WAVE] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t s = 0;
    for (uint32_t w = 0; w < blockDim.x / WAVE; ++w) s += s_part[w];
    Q.words[2] = (uint32_t)s;
    Q.words[3] = (uint32_t)(s >> 32);
  }
}

bool scan_batchable(const mtl_query& q) {
  if (q.kind != MTL_Q_SCAN || q.n_preds != 1 || q.n_rows == 0 || q.n_rows > 0xFFFFFFFFull)
    return false;
  const mtl_pred& p = q.preds[0];
  if (!p.column || !type_width(p.type) || p.type == MTL_HUGEINT || p.op < 0 || p.op > 4)
    return false;
  if (q.proj_col) {
    int pk = 0;
    if (proj_kind(q.proj_type, q.proj_op, &pk) != MTL_OK) return false;
  }
  return true;
}

bool scan_shareable(const mtl_query& a, const mtl_query& b) {
  return a.preds[0].column == b.preds[0].column && a.preds[0].type == b.preds[0].type &&
         a.proj_col == b.proj_col && (!a.proj_col || a.proj_type == b.proj_type);
}

size_t scan_batch_stage_bytes(uint32_t n) { return (size_t)n * sizeof(MQScanQ) + 256; }

int32_t scan_batch_enqueue(const ScanBatchItem* items, uint32_t n, uint32_t* words,
                           void* host_stage, mtl_stream* s) {
  if (n == 0) return MTL_OK;
  if (n > kMQMax) return set_error(MTL_ERR_INVALID_ARGUMENT, "more than %u queries", kMQMax);
  const mtl_query& q0 = *items[0].q;
  for (uint32_t i = 0; i < n; ++i)
    if (!scan_batchable(*items[i].q) || !scan_shareable(q0, *items[i].q))
      return set_error(MTL_ERR_INVALID_ARGUMENT, "query %u cannot join the shared scan", i);
  int pk = 0;
  if (q0.proj_col) MTL_TRY(proj_kind(q0.proj_type, q0.proj_op, &pk));
  Arena& ar = s->arena;
  // per-query scratch; the regions to clear (group sums, fingerprint sums)
  // are carved from one block so one clearing launch covers them
  uint64_t clear_words = 0;
  std::vector<uint32_t> tiles(n), groups(n);
  for (uint32_t i = 0; i < n; ++i) {
    tiles[i] = filter_tiles(items[i].q->n_rows);
    groups[i] = (uint32_t)ceil_div(tiles[i], 1u << kGroupShift);
    clear_words += (uint64_t)groups[i] * 2 * kFpStride;
  }
  void* clr;
  MTL_TRY(ar.alloc(clear_words * 4, &clr));
  MQScanQ* hq = (MQScanQ*)host_stage;
  uint32_t* cw = (uint32_t*)clr;
  for (uint32_t i = 0; i < n; ++i) {
    const mtl_query& q = *items[i].q;
    MQScanQ& Q = hq[i];
    Q = MQScanQ{};
    Q.pred = compile_pred(q.preds[0]);
    Q.n = q.n_rows;
    Q.tiles = tiles[i];
    Q.op = q.proj_op;
    Q.scalar = q.proj_scalar;
    Q.out_sel = items[i].out_sel;
    Q.out_proj = items[i].out_proj;
    void *bm, *cnt;
    MTL_TRY(ar.alloc((uint64_t)tiles[i] * FTILE / 8, &bm));
    MTL_TRY(ar.alloc((uint64_t)tiles[i] * 4, &cnt));
    Q.bm = (uint64_t*)bm;
    Q.counts = (uint32_t*)cnt;
    Q.fp = (uint64_t*)cw;  // 8-byte aligned: every carve is a multiple of 64 words
    cw += (uint64_t)groups[i] * 2 * kFpStride;
    Q.words = words + 4 * i;
  }
  void* dq;
  MTL_TRY(ar.alloc((size_t)n * sizeof(MQScanQ), &dq));
  hipStream_t st = s->stream;
  MTL_HIP(hipMemcpyAsync(dq, hq, (size_t)n * sizeof(MQScanQ), hipMemcpyHostToDevice, st));
  ZeroList zl;
  zl.add((uint32_t*)clr, clear_words);
  zl.add(words, (uint64_t)n * 4);
  MTL_TRY(launch_zero(zl, st));
  MQScanArgs a{};
  a.col = q0.preds[0].column;
  a.pcol = q0.proj_col;
  a.nq = n;
  a.q = (const MQScanQ*)dq;
  uint32_t tmax = 0;
  for (uint32_t i = 0; i < n; ++i) {
    tmax = std::max(tmax, tiles[i]);
    a.n_max = std::max<uint64_t>(a.n_max, items[i].q->n_rows);
  }
  const dim3 grid(tmax);
  {
    ProfScope ps(s, "k_mq_filter");
    switch (hq[0].pred.kind) {
      case PK_I32: hipLaunchKernelGGL((k_mq_filter<uint32_t, PK_I32>), grid, dim3(MQ_BLOCK), 0, st, a); break;
      case PK_F32: hipLaunchKernelGGL((k_mq_filter<uint32_t, PK_F32>), grid, dim3(MQ_BLOCK), 0, st, a); break;
      case PK_U8: hipLaunchKernelGGL((k_mq_filter<uint8_t, PK_U8>), grid, dim3(MQ_BLOCK), 0, st, a); break;
      case PK_I64: hipLaunchKernelGGL((k_mq_filter<uint64_t, PK_I64>), grid, dim3(MQ_BLOCK), 0, st, a); break;
      default: hipLaunchKernelGGL((k_mq_filter<uint64_t, PK_F64>), grid, dim3(MQ_BLOCK), 0, st, a); break;
    }
    MTL_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_mq_offsets, dim3(n), dim3(1024), 0, st, a.q);
  MTL_HIP(hipGetLastError());
  {
    ProfScope ps(s, "k_mq_expand");
    switch (pk) {
      case 1: hipLaunchKernelGGL(k_mq_expand<1>, grid, dim3(MQ_BLOCK), 0, st, a); break;
      case 2: hipLaunchKernelGGL(k_mq_expand<2>, grid, dim3(MQ_BLOCK), 0, st, a); break;
      case 3: hipLaunchKernelGGL(k_mq_expand<3>, grid, dim3(MQ_BLOCK), 0, st, a); break;
      case 4: hipLaunchKernelGGL(k_mq_expand<4>, grid, dim3(MQ_BLOCK), 0, st, a); break;
      default: hipLaunchKernelGGL(k_mq_expand<0>, grid, dim3(MQ_BLOCK), 0, st, a); break;
    }
    MTL_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_mq_fp_sum, dim3(n), dim3(1024), 0, st, a.q);
  MTL_HIP(hipGetLastError());
  return MTL_OK;
}

}  // namespace mtl

// ---------------------------------------------------------------------------
// Small scans of a batch in one persistent launch.  A scan of at most
// kSmallRows rows (any columns, up to MTL_MAX_PREDS predicates, an optional
// projection) is done by one block from start to end: per 4096-row tile the
// predicates, a block prefix of the passing rows (one barrier pair), the
// row ids and projected values at their output positions, and the result
// fingerprint; the block then writes the query's (count, overflow flag,
// fingerprint).  No kernel boundary or host step separates a query's
// filter from its compaction, and every small scan of the batch shares one
// launch: a fixed grid of blocks that take queries from a ticket until the
// batch's count is reached.  The launch sequence (descriptor upload from
// pinned memory, the kernel, the result download) is captured once per
// serving slot into a HIP graph and replayed for each batch.
// ---------------------------------------------------------------------------
namespace mtl {

struct SmallScanQ {
  PredC preds[MTL_MAX_PREDS];
  uint32_t n_preds;
  uint32_t pk;  // projection kind: 0 none, 1 int32, 2 int64, 3 float, 4 double
  uint64_t n;
  const void* pcol;
  int32_t op;
  int64_t scalar;
  uint32_t* out_sel;
  void* out_proj;
};

// The pinned (and device) block the graph copies: header, then descriptors.
struct SmallScanBlock {
  uint32_t nq;
  uint32_t ticket;  // uploaded as 0 with every batch
  uint32_t pad[2];
  SmallScanQ q[kSmallMax];
};

__device__ __forceinline__ uint64_t proj_store_rt(uint32_t pk, const ExpandProj& pj, uint64_t o,
                                                  uint64_t x, bool& ovf) {
  switch (pk) {
    case 1: return proj_store<1>(pj, o, x, ovf);
    case 2: return proj_store<2>(pj, o, x, ovf);
    case 3: return proj_store<3>(pj, o, x, ovf);
    case 4: return proj_store<4>(pj, o, x, ovf);
    default: return 0;
  }
}

// BLOCK threads; J 256-row steps per wave and tile (lane l of step j holds
// rows 4l..4l+3 of it): a tile is BLOCK * 4 * J rows.  Both are tuned
// (T_SMALL_BLOCK, T_SMALL_STEPS); every pair gives the same results.
template <int BLOCK, int J>
__global__ __launch_bounds__(BLOCK) void k_small_scan(SmallScanBlock* blk, uint32_t* res) {
  constexpr int W = BLOCK / WAVE;
  constexpr uint64_t TILE = (uint64_t)BLOCK * 4 * J;
  __shared__ uint32_t s_q;
  __shared__ uint32_t s_wtot[W];
  __shared__ uint64_t s_h[W];
  __shared__ uint32_t s_ovf;
  const int tid = threadIdx.x, wave = tid / WAVE, lane = tid % WAVE;
  const uint32_t nq = blk->nq;
  while (true) {
    if (tid == 0) s_q = atomicAdd(&blk->ticket, 1u);
    if (tid == 0) s_ovf = 0;
    __syncthreads();
    const uint32_t q = s_q;
    __syncthreads();
    if (q >= nq) break;  // block-uniform
    const SmallScanQ& Q = blk->q[q];
    const uint64_t n = Q.n;
    const uint32_t pk = Q.pk;
    const ExpandProj pj{Q.pcol, Q.op, Q.scalar, Q.out_proj, nullptr};
    const uint32_t pw = pk == 2 || pk == 4 ? 8u : 4u;
    uint32_t base = 0;
    bool ovf = false;
    uint64_t h = 0;
    for (uint64_t t0 = 0; t0 < n; t0 += TILE) {
      const uint64_t k0 = t0 + (uint64_t)wave * (J * 256) + (uint64_t)lane * 4;
      uint32_t bits = 0;
#pragma unroll
      for (int j = 0; j < J; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (k0 + (uint64_t)j * 256 + v < n) bits |= 1u << (4 * j + v);
      for (uint32_t p = 0; p < Q.n_preds; ++p) {
        const PredC& P = Q.preds[p];
        const uint32_t w = pred_width(P.kind);
        // all loads issued before any compare (rows past n read row n - 1,
        // in bounds, and are masked above)
        uint64_t x[J][4];
#pragma unroll
        for (int j = 0; j < J; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const uint64_t r = k0 + (uint64_t)j * 256 + v;
            x[j][v] = load_raw(P.col, w, r < n ? r : n - 1);
          }
#pragma unroll
        for (int j = 0; j < J; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v)
            if (!pred_pass(enc_raw(x[j][v], P.kind), P)) bits &= ~(1u << (4 * j + v));
      }
      // positions: wave prefix per step, then the tile's wave offsets
      uint32_t ej[J];
      uint32_t wsum = 0;
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const uint32_t c = (uint32_t)__popc((bits >> (4 * j)) & 0xFu);
        const uint32_t incl = wave_inclusive_u32(c);
        ej[j] = wsum + incl - c;
        wsum += (uint32_t)__builtin_amdgcn_readlane((int)incl, WAVE - 1);
      }
      if (lane == 0) s_wtot[wave] = wsum;
      __syncthreads();
      uint32_t woff = 0, ttot = 0;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        if (w < wave) woff += s_wtot[w];
        ttot += s_wtot[w];
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < J; ++j) {
        uint32_t k = 0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          if (!((bits >> (4 * j + v)) & 1u)) continue;
          const uint64_t r = k0 + (uint64_t)j * 256 + v;
          const uint64_t o = (uint64_t)base + woff + ej[j] + k++;
          Q.out_sel[o] = (uint32_t)r;
          const uint64_t x = pk ? load_raw(Q.pcol, pw, r) : 0ull;
          const uint64_t vb = proj_store_rt(pk, pj, o, x, ovf);
          h += fp_term(o, (uint32_t)r, pk ? vb : r);
        }
      }
      base += ttot;
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) h += __shfl_xor(h, d, WAVE);
    if (lane == 0) s_h[wave] = h;
    if (ovf) atomicOr(&s_ovf, 1u);
    __syncthreads();
    if (tid == 0) {
      uint64_t t = 0;
#pragma unroll
      for (int w = 0; w < W; ++w) t += s_h[w];
      res[4 * q] = base;
      res[4 * q + 1] = s_ovf;
      res[4 * q + 2] = (uint32_t)t;
      res[4 * q + 3] = (uint32_t)(t >> 32);
    }
    __syncthreads();
  }
}

size_t small_block_bytes() { return sizeof(SmallScanBlock); }

int32_t small_scan_launch(int32_t block, int32_t steps, uint32_t grid, void* d_blk,
                          uint32_t* d_res, hipStream_t st) {
  SmallScanBlock* b = (SmallScanBlock*)d_blk;
#define MTL_SMALL_CASE(B, S)                                                            \
  if (block == B && steps == S) {                                                      \
    hipLaunchKernelGGL((k_small_scan<B, S>), dim3(grid), dim3(B), 0, st, b, d_res);      \
    MTL_HIP(hipGetLastError());                                                        \
    return MTL_OK;                                                                     \
  }
  MTL_SMALL_CASE(128, 1) MTL_SMALL_CASE(128, 2) MTL_SMALL_CASE(128, 4)
  MTL_SMALL_CASE(256, 1) MTL_SMALL_CASE(256, 2) MTL_SMALL_CASE(256, 4)
  MTL_SMALL_CASE(512, 1) MTL_SMALL_CASE(512, 2) MTL_SMALL_CASE(512, 4)
  MTL_SMALL_CASE(1024, 1) MTL_SMALL_CASE(1024, 2) MTL_SMALL_CASE(1024, 4)
#undef MTL_SMALL_CASE
  return set_error(MTL_ERR_INVALID_ARGUMENT, "small scan block %d steps %d", block, steps);
}

bool small_scan_ok(const mtl_query& q) {
  if (q.kind != MTL_Q_SCAN || q.n_rows == 0 || q.n_rows > kSmallRows || q.n_preds == 0 ||
      q.n_preds > MTL_MAX_PREDS)
    return false;
  for (uint32_t i = 0; i < q.n_preds; ++i) {
    const mtl_pred& p = q.preds[i];
    if (!p.column || !type_width(p.type) || p.type == MTL_HUGEINT || p.op < 0 || p.op > 4)
      return false;
  }
  if (q.proj_col) {
    int pk = 0;
    if (proj_kind(q.proj_type, q.proj_op, &pk) != MTL_OK) return false;
  }
  return true;
}

struct SmallGraph {
  SmallScanBlock* h_blk = nullptr;  // pinned
  SmallScanBlock* d_blk = nullptr;
  uint32_t* h_res = nullptr;  // pinned, 4 words per query
  uint32_t* d_res = nullptr;
  hipGraphExec_t exec = nullptr;
};

static void small_graph_free(SmallGraph* g) {
  if (!g) return;
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->h_blk) hipHostFree(g->h_blk);
  if (g->h_res) hipHostFree(g->h_res);
  dev_free(g->d_blk, sizeof(SmallScanBlock));
  dev_free(g->d_res, kSmallMax * 16);
  delete g;
}

void small_graph_destroy(SmallGraph* g) { small_graph_free(g); }

int32_t small_graph_create(mtl_stream* s, SmallGraph** out) {
  auto* g = new SmallGraph();
  int32_t st = MTL_OK;
  if (hipHostMalloc((void**)&g->h_blk, sizeof(SmallScanBlock), hipHostMallocDefault) !=
          hipSuccess ||
      hipHostMalloc((void**)&g->h_res, kSmallMax * 16, hipHostMallocDefault) != hipSuccess)
    st = set_error(MTL_ERR_PROCESSING, "processing error: pinned small-scan staging");
  if (st == MTL_OK) st = dev_alloc(sizeof(SmallScanBlock), (void**)&g->d_blk, "small scans");
  if (st == MTL_OK) st = dev_alloc(kSmallMax * 16, (void**)&g->d_res, "small scans");
  if (st != MTL_OK) {
    small_graph_free(g);
    return st;
  }
  // capture: upload the block, run the persistent kernel, download the
  // results (the ticket is uploaded as 0, so no clearing node is needed)
  // with the tuned block shape at capture time; one block per query of the
  // largest batch
  hipStream_t cs = s->stream;
  hipGraph_t graph = nullptr;
  const uint32_t grid = kSmallMax;
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  int32_t lst = MTL_OK;
  if (e == hipSuccess) {
    hipMemcpyAsync(g->d_blk, g->h_blk, sizeof(SmallScanBlock), hipMemcpyHostToDevice, cs);
    lst = small_scan_launch(tuned(T_SMALL_BLOCK), tuned(T_SMALL_STEPS), grid, g->d_blk,
                            g->d_res, cs);
    hipMemcpyAsync(g->h_res, g->d_res, kSmallMax * 16, hipMemcpyDeviceToHost, cs);
    e = hipStreamEndCapture(cs, &graph);
  }
  if (lst != MTL_OK) {
    if (graph) hipGraphDestroy(graph);
    small_graph_free(g);
    return lst;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
  if (graph) hipGraphDestroy(graph);
  if (e != hipSuccess) {
    small_graph_free(g);
    return set_error(MTL_ERR_PROCESSING, "processing error: small-scan graph: %s",
                     hipGetErrorString(e));
  }
  *out = g;
  return MTL_OK;
}

int32_t small_scan_stage(void* h_blk, const SmallScanItem* items, uint32_t n) {
  if (n > kSmallMax) return set_error(MTL_ERR_INVALID_ARGUMENT, "more than %u small scans",
                                      kSmallMax);
  SmallScanBlock* b = (SmallScanBlock*)h_blk;
  b->nq = n;
  b->ticket = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const mtl_query& q = *items[i].q;
    if (!small_scan_ok(q))
      return set_error(MTL_ERR_INVALID_ARGUMENT, "query %u is not a small scan", i);
    SmallScanQ& Q = b->q[i];
    Q = SmallScanQ{};
    for (uint32_t p = 0; p < q.n_preds; ++p) Q.preds[p] = compile_pred(q.preds[p]);
    Q.n_preds = q.n_preds;
    Q.n = q.n_rows;
    int pk = 0;
    if (q.proj_col) MTL_TRY(proj_kind(q.proj_type, q.proj_op, &pk));
    Q.pk = (uint32_t)pk;
    Q.pcol = q.proj_col;
    Q.op = q.proj_op;
    Q.scalar = q.proj_scalar;
    Q.out_sel = items[i].out_sel;
    Q.out_proj = items[i].out_proj;
  }
  return MTL_OK;
}

int32_t small_scan_enqueue(SmallGraph* g, const SmallScanItem* items, uint32_t n,
                           mtl_stream* s) {
  if (n == 0) return MTL_OK;
  MTL_TRY(small_scan_stage(g->h_blk, items, n));
  MTL_HIP(hipGraphLaunch(g->exec, s->stream));
  return MTL_OK;
}

const uint32_t* small_scan_results(const SmallGraph* g) { return g->h_res; }

}  // namespace mtl
