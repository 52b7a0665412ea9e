// Block-level top-K over candidates keyed by up to four 64-bit words plus a
// 32-bit id, lexicographically smaller = better.  Used by the star pipeline
// and by the grouped-aggregate top-K.
//
// Reference: grouped_candidate_better (kernels.metal:1102-1118) compared
// (sum desc, secondary asc, row id asc) and topk_dense_grouped_rows_parallel_v
// (:1249-1390) ran one 128-thread group with K <= 20.  Here every block
// keeps a candidate buffer in LDS, admits only entries whose primary key
// does not lose to the current K-th best, and the last block to finish
// merges all block results; K <= TOPK_MAX_K.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mtl {
namespace dev {

constexpr int TOPK_BLOCK = 256;
constexpr int TOPK_CAP = 1024;          // LDS candidate buffer entries
constexpr int TOPK_ROUND = TOPK_BLOCK;  // entries admitted per round (1 per thread)
constexpr int TOPK_MAX_K = TOPK_CAP - 2 * TOPK_ROUND;  // 512

struct TopkLds {
  uint64_t k[4][TOPK_CAP];
  uint32_t id[TOPK_CAP];
  uint32_t n;
  uint64_t thr;  // admit entries with k[0] <= thr
};

__device__ __forceinline__ bool cand_less(const TopkLds& s, int a, int b) {
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    if (s.k[f][a] != s.k[f][b]) return s.k[f][a] < s.k[f][b];
  }
  return s.id[a] < s.id[b];
}

__device__ __forceinline__ void cand_swap(TopkLds& s, int a, int b) {
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    uint64_t t = s.k[f][a];
    s.k[f][a] = s.k[f][b];
    s.k[f][b] = t;
  }
  uint32_t t = s.id[a];
  s.id[a] = s.id[b];
  s.id[b] = t;
}

// Sort the first n entries ascending (best first); pads to a power of two.
// All threads of the block must call it.
__device__ void topk_sort(TopkLds& s, uint32_t n) {
  uint32_t p = 2;
  while (p < n) p <<= 1;
  for (uint32_t i = n + threadIdx.x; i < p; i += blockDim.x) {
#pragma unroll
    for (int f = 0; f < 4; ++f) s.k[f][i] = ~0ull;
    s.id[i] = 0xFFFFFFFFu;
  }
  __syncthreads();
  for (uint32_t size = 2; size <= p; size <<= 1) {
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (uint32_t t = threadIdx.x; t < p / 2; t += blockDim.x) {
        uint32_t i = 2 * t - (t & (stride - 1));
        uint32_t j = i + stride;
        bool up = (i & size) == 0;
        bool worse = cand_less(s, j, i);  // a[i] > a[j]
        if (worse == up) cand_swap(s, i, j);
      }
      __syncthreads();
    }
  }
}

// After a sort: keep the best k and tighten the admission threshold.
__device__ __forceinline__ void topk_truncate(TopkLds& s, uint32_t k) {
  if (threadIdx.x == 0) {
    if (s.n > k) s.n = k;
    if (s.n == k && k > 0) s.thr = s.k[0][k - 1];
  }
  __syncthreads();
}

__device__ __forceinline__ void topk_init(TopkLds& s) {
  if (threadIdx.x == 0) {
    s.n = 0;
    s.thr = ~0ull;
  }
  __syncthreads();
}

// Append a candidate (caller checked k0 <= thr).
__device__ __forceinline__ void topk_push(TopkLds& s, uint64_t k0, uint64_t k1,
                                          uint64_t k2, uint64_t k3, uint32_t id) {
  uint32_t pos = atomicAdd(&s.n, 1u);
  s.k[0][pos] = k0;
  s.k[1][pos] = k1;
  s.k[2][pos] = k2;
  s.k[3][pos] = k3;
  s.id[pos] = id;
}

// Called by all threads after each admission round (<= TOPK_ROUND pushes).
// Compacts as soon as the buffer holds more than max(64, k) entries, so the
// first sort (which sets the threshold) happens after the first round;
// after it most entries fail the threshold test and are never pushed.
// Bound: after a compaction n <= k, a round adds <= TOPK_ROUND, so n never
// exceeds max(64, k) + TOPK_ROUND <= TOPK_CAP.
__device__ __forceinline__ void topk_maybe_compact(TopkLds& s, uint32_t k) {
  __syncthreads();
  const uint32_t limit = k > 64u ? k : 64u;
  if (s.n > limit) {
    uint32_t n = s.n;
    __syncthreads();
    topk_sort(s, n);
    topk_truncate(s, k);
  }
}

// Sort the block's candidates, publish its best k to slot blockIdx.x of the
// global candidate arrays (ck[f][b*k + i], cid, cn[b]), and let the last
// block to finish merge all published candidates into s.  Returns true
// only in that last block, with s holding the final top-k in order.  The
// merge reads every block's slots in parallel (flat index b*k + i).
__device__ __forceinline__ bool topk_publish_elect_merge(TopkLds& s, uint64_t* const* ck,
                                                         uint32_t* cid, uint32_t* cn,
                                                         uint32_t* done, uint32_t k) {
  __shared__ bool s_last;
  __syncthreads();
  uint32_t n = s.n;
  __syncthreads();
  topk_sort(s, n);
  topk_truncate(s, k);
  const uint32_t kb = s.n;
  for (uint32_t i = threadIdx.x; i < kb; i += blockDim.x) {
    uint64_t o = (uint64_t)blockIdx.x * k + i;
#pragma unroll
    for (int f = 0; f < 4; ++f) ck[f][o] = s.k[f][i];
    cid[o] = s.id[i];
  }
  if (threadIdx.x == 0) cn[blockIdx.x] = kb;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) s_last = atomicAdd(done, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
  topk_init(s);
  // Each thread first loads up to E of the published candidates into
  // registers (all loads in flight together), then the admission rounds run
  // from registers: the merge costs one memory round trip per chunk of
  // TOPK_ROUND * E candidates instead of one per round.
  constexpr int E = 8;
  const uint32_t total = gridDim.x * k;
  for (uint32_t chunk = 0; chunk < total; chunk += TOPK_ROUND * E) {
    uint64_t c0[E], c1[E], c2[E], c3[E];
    uint32_t ci[E];
    bool ok[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      uint32_t i = chunk + e * TOPK_ROUND + threadIdx.x;
      uint32_t b = i / k, j = i - b * k;
      ok[e] = i < total && j < __hip_atomic_load(&cn[b < gridDim.x ? b : 0], __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
      uint32_t ii = ok[e] ? i : 0;
      c0[e] = ck[0][ii];
      c1[e] = ck[1][ii];
      c2[e] = ck[2][ii];
      c3[e] = ck[3][ii];
      ci[e] = cid[ii];
    }
    for (int e = 0; e < E; ++e) {
      if (chunk + e * TOPK_ROUND >= total) break;  // uniform
      if (ok[e] && c0[e] <= s.thr) topk_push(s, c0[e], c1[e], c2[e], c3[e], ci[e]);
      topk_maybe_compact(s, k);
    }
  }
  __syncthreads();
  n = s.n;
  __syncthreads();
  topk_sort(s, n);
  topk_truncate(s, k);
  return true;
}

}  // namespace dev
}  // namespace mtl
