// Top-K over candidates keyed by four 64-bit words plus a 32-bit id,
// lexicographically smaller = better (ids are unique, so the order is
// total).  Used by the star pipeline and by the grouped-aggregate top-K.
//
// Reference: grouped_candidate_better (kernels.metal:1102-1118) compared
// (sum desc, secondary asc, row id asc) and topk_dense_grouped_rows_parallel_v
// (:1249-1390) ran one 128-thread group with K <= 20.
//
// Here each wave keeps its own sorted list of the best K candidates in
// registers: position p = r*64 + lane lives in row r of that lane (R rows,
// K <= 64*R).  A batch of 64 candidates (one per lane) is first tested
// against the K-th key word 0 of the wave and a block-shared bound (one
// compare per lane, one ballot); the few that pass are inserted one at a
// time: rank = popcount of a ballot of "entry better than candidate", then
// the tail shifts down one position with DPP lane moves.  The wave lists of
// a block, then the block lists, are merged by rank in LDS (below), level by
// level, with one binary search per entry instead of one insertion per
// candidate.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mtl {
namespace dev {

constexpr int TOPK_MAX_K = 512;

// Phase-1 block shape for lists of R rows: 16 waves while the per-wave
// lists are short, 4 waves for K > 128 (LDS holds two buffers of W lists).
template <int R>
struct TopkShape {
  static constexpr int NT = R <= 2 ? 1024 : 256;
  static constexpr int W = NT / 64;
  static constexpr int CAP = 64 * R;
};

__device__ __forceinline__ uint64_t bcast64(uint64_t v, int src) {
  uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)v, src);
  uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), src);
  return ((uint64_t)hi << 32) | lo;
}
// Lane i receives lane i-1's value (lane 0 keeps its own): one DPP move
// per 32 bits (wave_shr:1), no LDS crossbar round trip.
__device__ __forceinline__ uint32_t shup32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x138, 0xF, 0xF, false);
}
__device__ __forceinline__ uint64_t shup64(uint64_t v) {
  return ((uint64_t)shup32((uint32_t)(v >> 32)) << 32) | shup32((uint32_t)v);
}

template <int R>
struct WaveTopk {
  uint64_t k[R][4];
  uint32_t id[R];
  uint32_t n;     // valid entries (wave-uniform)
  uint32_t K;
  uint64_t thr0;  // key word 0 of the K-th entry once full, else all ones

  __device__ __forceinline__ void init(uint32_t kk) {
    n = 0;
    K = kk;
    thr0 = ~0ull;
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
      for (int f = 0; f < 4; ++f) k[r][f] = ~0ull;
      id[r] = 0xFFFFFFFFu;
    }
  }

  // entry (row r of this lane) strictly better than c
  __device__ __forceinline__ bool better(int r, const uint64_t (&c)[4], uint32_t cid) const {
#pragma unroll
    for (int f = 0; f < 4; ++f)
      if (k[r][f] != c[f]) return k[r][f] < c[f];
    return id[r] < cid;
  }

  // Insert one wave-uniform candidate.
  __device__ __forceinline__ void insert(const uint64_t (&c)[4], uint32_t cid) {
    const int lane = __lane_id();
    uint32_t pos = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const bool b = (uint32_t)(r * 64 + lane) < n && better(r, c, cid);
      pos += (uint32_t)__popcll(__ballot(b));
    }
    if (pos >= K) return;
#pragma unroll
    for (int r = R - 1; r >= 0; --r) {
      if ((uint32_t)(r * 64) > n) continue;  // row holds nothing and receives nothing
      uint64_t pk[4];
      uint32_t pid;
#pragma unroll
      for (int f = 0; f < 4; ++f) pk[f] = shup64(k[r][f]);
      pid = shup32(id[r]);
      if (r > 0) {
        uint64_t ck[4];
#pragma unroll
        for (int f = 0; f < 4; ++f) ck[f] = bcast64(k[r - 1][f], 63);
        uint32_t cidr = __builtin_amdgcn_readlane((int)id[r - 1], 63);
        if (lane == 0) {
#pragma unroll
          for (int f = 0; f < 4; ++f) pk[f] = ck[f];
          pid = cidr;
        }
      }
      const uint32_t p = (uint32_t)(r * 64 + lane);
      if (p > pos) {
#pragma unroll
        for (int f = 0; f < 4; ++f) k[r][f] = pk[f];
        id[r] = pid;
      } else if (p == pos) {
#pragma unroll
        for (int f = 0; f < 4; ++f) k[r][f] = c[f];
        id[r] = cid;
      }
    }
    if (n < K) ++n;
    if (n == K) {
      const int rr = (int)((K - 1) / 64), ll = (int)((K - 1) % 64);
      uint64_t t = 0;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (r == rr) t = bcast64(k[r][0], ll);
      thr0 = t;
    }
  }

  // Each lane offers at most one candidate (ok = it passed the caller's
  // k0 <= thr0 test, keys complete).  Wave-uniform control flow required.
  __device__ __forceinline__ void offer(bool ok, const uint64_t (&c)[4], uint32_t cid) {
    uint64_t mask = __ballot(ok);
    while (mask) {
      const int src = __builtin_amdgcn_readfirstlane(__ffsll((unsigned long long)mask) - 1);
      uint64_t bc[4];
#pragma unroll
      for (int f = 0; f < 4; ++f) bc[f] = bcast64(c[f], src);
      const uint32_t bid = __builtin_amdgcn_readlane((int)cid, src);
      insert(bc, bid);
      mask &= mask - 1;
      if (n == K) mask &= __ballot(c[0] <= thr0);
    }
  }
};

// ---------------------------------------------------------------------------
// Merging sorted lists: merge by rank in LDS.
//
// Every wave list entry is staged in global memory with its full key
// (stage.ck[f][g], stage.cid[g]); merges move only (key word 0, g).  Two
// sorted lists A and B merge in one parallel step: entry i of A lands at
// i + |{b in B : b < A[i]}| (a binary search in B), and symmetrically; the
// keys are distinct (ids are unique), so the positions are a permutation and
// entries landing at K or beyond are dropped.  L lists merge pairwise in
// ceil(log2 L) levels with one barrier each.  A tie on word 0 reads the
// full keys from the stage (rare: equal aggregates).
// ---------------------------------------------------------------------------
struct TopkStage {
  uint64_t* ck[4];
  uint32_t* cid;
};

__device__ __forceinline__ bool tk_less(uint64_t ak, uint32_t ag, uint64_t bk, uint32_t bg,
                                        const TopkStage& st) {
  if (ak != bk) return ak < bk;
#pragma unroll
  for (int f = 1; f < 4; ++f) {
    const uint64_t x = st.ck[f][ag], y = st.ck[f][bg];
    if (x != y) return x < y;
  }
  return st.cid[ag] < st.cid[bg];
}

// Lists li < L of length n[li] <= K at stride K in buffer `cur` (k0, g, n);
// returns the buffer holding the merged list (list 0).  Every thread of
// the block calls it after a barrier that published the input.
template <int NT>
__device__ __forceinline__ int tk_tree_merge(uint64_t* const (&k0)[2], uint32_t* const (&g)[2],
                                             uint32_t* const (&n)[2], uint32_t L, uint32_t K,
                                             const TopkStage& st) {
  int cur = 0;
  while (L > 1) {  // block-uniform
    const uint32_t Lo = (L + 1) >> 1;
    const uint64_t* ak = k0[cur];
    const uint32_t* ag = g[cur];
    const uint32_t* an = n[cur];
    uint64_t* bk = k0[cur ^ 1];
    uint32_t* bg = g[cur ^ 1];
    uint32_t* bn = n[cur ^ 1];
    for (uint32_t t = threadIdx.x; t < L * K; t += NT) {
      const uint32_t li = t / K, i = t - li * K;
      if (i >= an[li]) continue;
      const uint64_t xk = ak[t];
      const uint32_t xg = ag[t];
      uint32_t pos = i;
      const uint32_t pl = li ^ 1u;
      if (pl < L) {
        const uint32_t base = pl * K;
        uint32_t lo = 0, hi = an[pl];
        while (lo < hi) {
          const uint32_t mid = (lo + hi) >> 1;
          if (tk_less(ak[base + mid], ag[base + mid], xk, xg, st)) lo = mid + 1;
          else hi = mid;
        }
        pos += lo;
      }
      if (pos < K) {
        bk[(li >> 1) * K + pos] = xk;
        bg[(li >> 1) * K + pos] = xg;
      }
    }
    for (uint32_t q = threadIdx.x; q < Lo; q += NT) {
      const uint32_t c = an[2 * q] + (2 * q + 1 < L ? an[2 * q + 1] : 0u);
      bn[q] = c < K ? c : K;
    }
    __syncthreads();
    cur ^= 1;
    L = Lo;
  }
  return cur;
}

// A set of lists (k0, g) at stride K with their lengths (global memory).
struct TopkLists {
  uint64_t* k0;
  uint32_t* g;
  uint32_t* n;
};

// LDS of a phase-1 block: one list per wave (W waves, K <= CAP per list).
template <int W, int CAP>
struct TopkBlockLds {
  uint64_t k0[2][W * CAP];
  uint32_t g[2][W * CAP];
  uint32_t n[2][W];
  uint64_t thr;  // block-shared bound on key word 0 (min of full waves' K-th)
};

template <int W, int CAP>
__device__ __forceinline__ void topk_block_init(TopkBlockLds<W, CAP>& s) {
  if (threadIdx.x == 0) s.thr = ~0ull;
  __syncthreads();
}

// Seed the block bound before any insertion (K <= 128).  Every lane passes
// the smallest and second-smallest key word 0 among the candidates it holds
// (lmin1 <= lmin2, all ones when absent).  For K <= 64 a wave sorts its 64
// lane minima (bitonic network over the lanes) and takes the K-th smallest:
// K lanes hold a candidate at or below it.  For K <= 128 it sorts the lanes'
// second minima and takes the ceil(K/2)-th smallest: ceil(K/2) lanes hold two
// candidates at or below it.  Either way no candidate above the bound can
// reach the top K, so the block bound is lowered to it.  With the bound known
// from the start, only a few candidates per block are ever inserted, instead
// of every candidate that beats a still-filling list.  Every thread of the
// block calls it (it ends with a barrier).
template <int W, int CAP>
__device__ __forceinline__ void topk_seed_thr(uint64_t lmin1, uint64_t lmin2, uint32_t K,
                                              TopkBlockLds<W, CAP>& s) {
  const int lane = __lane_id();
  uint64_t v = K <= 64 ? lmin1 : lmin2;
  const int at = K <= 64 ? (int)K - 1 : (int)((K + 1) / 2) - 1;
#pragma unroll
  for (int kk = 2; kk <= 64; kk <<= 1)
#pragma unroll
    for (int jj = kk >> 1; jj > 0; jj >>= 1) {
      const uint64_t o =
          ((uint64_t)(uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), jj, 64) << 32) |
          (uint32_t)__shfl_xor((int)(uint32_t)v, jj, 64);
      const bool asc = (lane & kk) == 0, lower = (lane & jj) == 0;
      v = (asc == lower) ? (o < v ? o : v) : (o > v ? o : v);
    }
  if (K >= 1 && K <= 128) {
    const uint64_t t = bcast64(v, at);
    if (lane == 0 && t != ~0ull && t < s.thr)
      atomicMin((unsigned long long*)&s.thr, (unsigned long long)t);
  }
  __syncthreads();
}

// Seed the block bound with the exact K-th smallest key word 0 among the
// candidates the block holds (v[e] where valid[e]): since those are K keys
// no worse than it, nothing above it can reach the top K.  A tighter bound
// than topk_seed_thr's (lane minima), for scans of many candidates per
// block, where each candidate that passes a loose bound costs a wave-serial
// insertion.  MSD radix select, 8 digits of 8 bits: per digit a 256-bin LDS
// histogram of the candidates still matching the chosen prefix, and one
// wave finds the digit that holds the rank.  Every thread of the block calls
// it (it ends with a barrier).
struct TopkSelectLds {
  uint32_t hist[256];
  uint32_t digit, rank, total;
};

template <int NT, int GE, int W, int CAP>
__device__ __forceinline__ void topk_seed_select(const uint64_t (&v)[GE], const bool (&valid)[GE],
                                                 uint32_t K, TopkSelectLds& q,
                                                 TopkBlockLds<W, CAP>& s) {
  const int tid = threadIdx.x, lane = __lane_id();
  uint64_t prefix = 0;
  uint32_t rank = K;  // 1-based rank still to find among the matching keys
  for (int d = 7; d >= 0; --d) {
    const int sh = d * 8;
    const uint64_t hi_mask = d == 7 ? 0ull : (~0ull << (sh + 8));
    for (int i = tid; i < 256; i += NT) q.hist[i] = 0;
    __syncthreads();
#pragma unroll
    for (int e = 0; e < GE; ++e)
      if (valid[e] && (v[e] & hi_mask) == prefix) atomicAdd(&q.hist[(v[e] >> sh) & 255u], 1u);
    __syncthreads();
    if (tid < 64) {
      // lane l holds bins 4l .. 4l+3: inclusive prefix over the lanes
      uint32_t c[4], sum = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        c[b] = q.hist[4 * lane + b];
        sum += c[b];
      }
      uint32_t incl = sum;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(incl, off, 64);
        if (lane >= off) incl += o;
      }
      if (lane == 63) q.total = incl;
      uint32_t before = incl - sum;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        if (before < rank && rank <= before + c[b]) {
          q.digit = (uint32_t)(4 * lane + b);
          q.rank = rank - before;
        }
        before += c[b];
      }
    }
    __syncthreads();
    if (q.total < rank) {  // fewer than K candidates: no bound (block-uniform)
      __syncthreads();
      return;
    }
    prefix |= (uint64_t)q.digit << sh;
    rank = q.rank;
    __syncthreads();  // q is rewritten by the next digit
  }
  if (tid == 0 && prefix < s.thr) atomicMin((unsigned long long*)&s.thr, (unsigned long long)prefix);
  __syncthreads();
}

// Running smallest and second-smallest of a lane's candidate keys.
__device__ __forceinline__ void topk_min2(uint64_t k0, uint64_t& m1, uint64_t& m2) {
  if (k0 < m1) {
    m2 = m1;
    m1 = k0;
  } else if (k0 < m2) {
    m2 = k0;
  }
}

// After a batch: a wave with a full list lowers the block bound.
template <int R, int W, int CAP>
__device__ __forceinline__ void topk_share_thr(const WaveTopk<R>& wl, TopkBlockLds<W, CAP>& s) {
  if (wl.n == wl.K && wl.K > 0 && __lane_id() == 0 && wl.thr0 < s.thr)
    atomicMin((unsigned long long*)&s.thr, (unsigned long long)wl.thr0);
}

// End of phase 1: stage every wave's entries, merge the block's wave lists
// in LDS and write the block's list to `out` (list blockIdx.x).  Every
// thread of the block calls it.
template <int NT, int R, int W, int CAP>
__device__ __forceinline__ void topk_block_finish(const WaveTopk<R>& wl, TopkBlockLds<W, CAP>& s,
                                                  const TopkStage& st, const TopkLists& out,
                                                  uint32_t bid) {
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const uint32_t K = wl.K;
  const uint32_t gbase = (bid * W + wave) * K;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t i = r * 64 + lane;
    if (i < wl.n) {
      const uint32_t gi = gbase + i;
#pragma unroll
      for (int f = 0; f < 4; ++f) st.ck[f][gi] = wl.k[r][f];
      st.cid[gi] = wl.id[r];
      s.k0[0][wave * K + i] = wl.k[r][0];
      s.g[0][wave * K + i] = gi;
    }
  }
  if (lane == 0) s.n[0][wave] = wl.n;
  __syncthreads();  // stage and LDS lists visible to the block
  uint64_t* const k0p[2] = {s.k0[0], s.k0[1]};
  uint32_t* const gp[2] = {s.g[0], s.g[1]};
  uint32_t* const np[2] = {s.n[0], s.n[1]};
  const int cur = tk_tree_merge<NT>(k0p, gp, np, W, K, st);
  const uint32_t nf = s.n[cur][0];
  for (uint32_t i = threadIdx.x; i < nf; i += NT) {
    out.k0[(uint64_t)bid * K + i] = s.k0[cur][i];
    out.g[(uint64_t)bid * K + i] = s.g[cur][i];
  }
  if (threadIdx.x == 0) out.n[bid] = nf;
}

// The same for a kernel whose block index is its list index.
template <int NT, int R, int W, int CAP>
__device__ __forceinline__ void topk_block_finish(const WaveTopk<R>& wl, TopkBlockLds<W, CAP>& s,
                                                  const TopkStage& st, const TopkLists& out) {
  topk_block_finish<NT>(wl, s, st, out, (uint32_t)blockIdx.x);
}

// Merge kernel body: block b merges lists [b*F, b*F + F) of `in` (L lists)
// in LDS; returns the LDS buffer holding the block's merged list.
constexpr int TOPK_MCAP = 2048;  // LDS entries per merge buffer (F * K)
constexpr int TOPK_MFAN = 256;   // max lists per merge block
constexpr int TOPK_MERGE_THREADS = 1024;

struct TopkMergeLds {
  uint64_t k0[2][TOPK_MCAP];
  uint32_t g[2][TOPK_MCAP];
  uint32_t n[2][TOPK_MFAN];
};

__host__ __device__ __forceinline__ uint32_t topk_fan(uint32_t K) {
  uint32_t f = (uint32_t)TOPK_MCAP / K;
  return f < (uint32_t)TOPK_MFAN ? f : (uint32_t)TOPK_MFAN;
}

__device__ __forceinline__ int topk_merge_lists(TopkMergeLds& s, const TopkLists& in, uint32_t L,
                                                uint32_t K, const TopkStage& st,
                                                uint32_t bid = 0xFFFFFFFFu) {
  const uint32_t F = topk_fan(K);
  const uint32_t first = (bid == 0xFFFFFFFFu ? (uint32_t)blockIdx.x : bid) * F;
  const uint32_t nl = L - first < F ? L - first : F;
  for (uint32_t t = threadIdx.x; t < nl * K; t += TOPK_MERGE_THREADS) {
    const uint64_t src = (uint64_t)first * K + t;  // entries past a list's length are unused
    s.k0[0][t] = in.k0[src];
    s.g[0][t] = in.g[src];
  }
  for (uint32_t li = threadIdx.x; li < nl; li += TOPK_MERGE_THREADS) s.n[0][li] = in.n[first + li];
  __syncthreads();
  uint64_t* const k0p[2] = {s.k0[0], s.k0[1]};
  uint32_t* const gp[2] = {s.g[0], s.g[1]};
  uint32_t* const np[2] = {s.n[0], s.n[1]};
  return tk_tree_merge<TOPK_MERGE_THREADS>(k0p, gp, np, nl, K, st);
}

__device__ __forceinline__ void topk_write_list(const TopkMergeLds& s, int cur, const TopkLists& out,
                                                uint32_t K, uint32_t bid = 0xFFFFFFFFu) {
  const uint32_t b = bid == 0xFFFFFFFFu ? (uint32_t)blockIdx.x : bid;
  const uint32_t nf = s.n[cur][0];
  for (uint32_t i = threadIdx.x; i < nf; i += TOPK_MERGE_THREADS) {
    out.k0[(uint64_t)b * K + i] = s.k0[cur][i];
    out.g[(uint64_t)b * K + i] = s.g[cur][i];
  }
  if (threadIdx.x == 0) out.n[b] = nf;
}

}  // namespace dev
}  // namespace mtl
