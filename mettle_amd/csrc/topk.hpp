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
// against the K-th key word 0 (one compare per lane, one ballot); the few
// that pass are inserted one at a time: rank = popcount of a ballot of
// "entry better than candidate", then the tail shifts down one position
// with lane shuffles.  No LDS and no block barrier on this path.  The waves
// of a block then merge through LDS into wave 0, each block publishes its
// list, and a one-block kernel merges the published lists the same way.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mtl {
namespace dev {

constexpr int TOPK_BLOCK = 256;
constexpr int TOPK_WAVES = TOPK_BLOCK / 64;
constexpr int TOPK_MAX_K = 512;

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

// LDS area for merging the waves of a block into wave 0.
template <int R>
struct TopkMergeLds {
  uint64_t k[TOPK_WAVES - 1][4][R * 64];
  uint32_t id[TOPK_WAVES - 1][R * 64];
  uint32_t n[TOPK_WAVES - 1];
};

// Merge the lists of all waves of the block into wave 0's list.  Every
// thread of the block must call it; afterwards only wave 0's list is
// meaningful.
template <int R>
__device__ __forceinline__ void topk_block_merge(WaveTopk<R>& wl, TopkMergeLds<R>& m) {
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (wave > 0) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
      for (int f = 0; f < 4; ++f) m.k[wave - 1][f][r * 64 + lane] = wl.k[r][f];
      m.id[wave - 1][r * 64 + lane] = wl.id[r];
    }
    if (lane == 0) m.n[wave - 1] = wl.n;
  }
  __syncthreads();
  if (wave == 0) {
    for (int w = 0; w < TOPK_WAVES - 1; ++w) {
      const uint32_t nw = m.n[w];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if ((uint32_t)(r * 64) >= nw) break;
        const uint32_t p = r * 64 + lane;
        uint64_t c[4];
#pragma unroll
        for (int f = 0; f < 4; ++f) c[f] = m.k[w][f][p];
        wl.offer(p < nw && c[0] <= wl.thr0, c, m.id[w][p]);
      }
    }
  }
}

// Published per-block lists: ck[f][b*K + i], cid[b*K + i], cn[b].
struct TopkPublished {
  uint64_t* ck[4];
  uint32_t* cid;
  uint32_t* cn;
};

// Wave 0 of block b writes its (merged) list.
template <int R>
__device__ __forceinline__ void topk_publish(const WaveTopk<R>& wl, const TopkPublished& p,
                                             uint32_t b) {
  if (threadIdx.x >= 64) return;
  const int lane = threadIdx.x;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t i = r * 64 + lane;
    if (i < wl.n) {
      const uint64_t o = (uint64_t)b * wl.K + i;
#pragma unroll
      for (int f = 0; f < 4; ++f) p.ck[f][o] = wl.k[r][f];
      p.cid[o] = wl.id[r];
    }
  }
  if (lane == 0) p.cn[b] = wl.n;
}

// One block: merge nb published lists into wave 0's list (sorted, best
// first).  Waves take the published entries round robin in batches of 64,
// with the loads of E batches in flight together.
template <int R>
__device__ __forceinline__ void topk_merge_published(WaveTopk<R>& wl, TopkMergeLds<R>& m,
                                                     const TopkPublished& p, uint32_t nb,
                                                     uint32_t K) {
  wl.init(K);
  const uint32_t total = nb * K;
  // every load is issued unconditionally (entries past a block's count are
  // read but discarded), so the counts do not delay the candidate loads
  constexpr int E = R <= 2 ? 16 : 4;
  for (uint32_t chunk = 0; chunk < total; chunk += TOPK_BLOCK * E) {
    uint64_t c[E][4];
    uint32_t ci[E], cnb[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const uint32_t i = chunk + e * TOPK_BLOCK + threadIdx.x;
      const uint32_t ii = i < total ? i : 0;
      cnb[e] = p.cn[ii / K];
#pragma unroll
      for (int f = 0; f < 4; ++f) c[e][f] = p.ck[f][ii];
      ci[e] = p.cid[ii];
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const uint32_t i = chunk + e * TOPK_BLOCK + threadIdx.x;
      const bool ok = i < total && i % K < cnb[e];
      wl.offer(ok && c[e][0] <= wl.thr0, c[e], ci[e]);
    }
  }
  topk_block_merge(wl, m);
}

// Broadcast wave 0's final list to LDS so any thread can read entry i.
template <int R>
__device__ __forceinline__ void topk_to_lds(const WaveTopk<R>& wl, uint32_t* ids, uint32_t* n) {
  if (threadIdx.x < 64) {
#pragma unroll
    for (int r = 0; r < R; ++r) ids[r * 64 + threadIdx.x] = wl.id[r];
    if (threadIdx.x == 0) *n = wl.n;
  }
  __syncthreads();
}

}  // namespace dev
}  // namespace mtl
