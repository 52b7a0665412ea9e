// Serving loop: queries arrive (optionally at scheduled times), wait in
// arrival order, are admitted while their device scratch fits the budget,
// and run concurrently on n_slots streams, each with its own worker thread,
// arena and result buffers.  A query that finishes frees its slot and its
// budget charge, and the next admissible query starts at once (continuous
// admission, not a batch barrier).  With n_slots == 1 it is the serial
// one-query-at-a-time dispatch the reference has (materialize_query,
// processing.rs:103-143, called once per CALL inside bind).
//
// Scan queries leave their rows on the device (the slot's selection and
// projection buffers); the host gets the row count and an order-sensitive
// fingerprint of the (row id, value) pairs, computed on the device, so the
// result of a query can be compared alone and in a mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <thread>
#include <unordered_map>

#include "device.hpp"
#include "filter.hpp"
#include "mtl_internal.hpp"

namespace mtl {
using namespace dev;

// Sum of the per-group fingerprint sums (stride apart) into *out.
__global__ void k_fingerprint_sum(const uint64_t* partial, uint32_t n, uint32_t stride,
                                  uint64_t* out) {
  __shared__ uint64_t s_part[16];
  uint64_t t = 0;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) t += partial[(uint64_t)i * stride];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, WAVE);
  if (threadIdx.x % WAVE == 0) s_part[threadIdx.x / WAVE] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t s = 0;
    for (uint32_t w = 0; w < blockDim.x / WAVE; ++w) s += s_part[w];
    *out = s;
  }
}

static int64_t now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

uint64_t query_scratch(const mtl_query& q) {
  if (q.kind == MTL_Q_SCAN) {
    const uint64_t tiles = ceil_div(q.n_rows ? q.n_rows : 1, kTileRows);
    const uint64_t pw = q.proj_col ? (uint64_t)std::max(type_width(q.proj_type), 4) : 0;
    const uint64_t groups = ceil_div(tiles, 1u << kGroupShift);
    // selection + projection, pass bits and tile counts, group sums and
    // fingerprint sums (a cache line each per group), clearing words
    return q.n_rows * (4 + pw) + tiles * (kTileRows / 8 + 4) + groups * 512 + (1u << 16);
  }
  const mtl_star_desc& d = q.star;
  const uint64_t pspan = key_off(d.p_key_max, d.p_key_min) + 1;
  const uint64_t dspan = d.d_rows ? key_off(d.d_key_max, d.d_key_min) + 1 : 0;
  const uint64_t ftiles = ceil_div(d.f_rows ? d.f_rows : 1, kTileRows);
  return d.p_rows * 20 + ftiles * kTileRows * 8 + pspan / 8 * 2 + dspan / 4 + d.d_rows / 4 +
         (uint64_t)d.k * 1024 * 64 + (1u << 20);
}

struct Req {
  mtl_query q;
  uint64_t need = 0;
  int64_t not_before = 0;
  bool done = false;
  mtl_query_result res{};
};

struct Slot {
  mtl_stream* s = nullptr;
  uint32_t* sel = nullptr;
  void* proj = nullptr;
  uint64_t cap = 0;  // rows the buffers hold
  uint32_t* words = nullptr;  // [0] count, [1] overflow flag, [2..3] fingerprint
  void* host = nullptr;       // pinned copy of words[0..3]
};

static int32_t slot_reserve(Slot& sl, uint64_t rows) {
  if (sl.sel && rows <= sl.cap) return MTL_OK;
  MTL_HIP(hipStreamSynchronize(sl.s->stream));
  if (sl.sel) MTL_HIP(hipFree(sl.sel));
  if (sl.proj) MTL_HIP(hipFree(sl.proj));
  sl.sel = nullptr;
  sl.proj = nullptr;
  sl.cap = 0;
  const uint64_t cap = std::max<uint64_t>(rows, 1 << 16);
  MTL_HIP(hipMalloc((void**)&sl.sel, cap * 4));
  MTL_HIP(hipMalloc(&sl.proj, cap * 8));
  sl.cap = cap;
  return MTL_OK;
}

// One clearing launch (the filter's, with the slot's words added), the
// filter pass, the expansion with the projection and the fingerprint fused,
// the fingerprint sum, and one 16-byte copy of (count, overflow, fingerprint).
static int32_t run_scan(Slot& sl, const mtl_query& q, mtl_query_result& r) {
  MTL_TRY(slot_reserve(sl, q.n_rows));
  hipStream_t st = sl.s->stream;
  ScanExtras ex;
  ex.zero.add(sl.words, 4);
  ex.fingerprint = true;
  CompactProj pj{q.proj_col, q.proj_type, q.proj_op, q.proj_scalar, sl.proj, sl.words + 1};
  MTL_TRY(filter_project_ex(q.preds, q.n_preds, q.n_rows, q.proj_col ? &pj : nullptr, sl.sel,
                            sl.words, &ex, sl.s));
  hipLaunchKernelGGL(k_fingerprint_sum, dim3(1), dim3(1024), 0, st, ex.fp_groups, ex.fp_n,
                     (uint32_t)kFpStride, (uint64_t*)(sl.words + 2));
  MTL_HIP(hipGetLastError());
  MTL_HIP(hipMemcpyAsync(sl.host, sl.words, 16, hipMemcpyDeviceToHost, st));
  MTL_HIP(hipStreamSynchronize(st));
  const uint32_t* hw = (const uint32_t*)sl.host;
  if (hw[0] > q.n_rows) return set_error(MTL_ERR_PROCESSING, "device scan failed");
  r.count = hw[0];
  r.fingerprint = *(const uint64_t*)(hw + 2);
  if (hw[1]) return set_error(MTL_ERR_OVERFLOW, "Overflow in projection");
  return MTL_OK;
}

}  // namespace mtl

struct mtl_server {
  std::vector<mtl::Slot> slots;
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> queue;  // tickets in submission order
  std::unordered_map<uint64_t, mtl::Req> reqs;
  uint64_t next_ticket = 1;
  uint64_t budget = 0, inflight = 0;
  bool stop = false;
};

namespace mtl {

static void worker(mtl_server* srv, size_t idx) {
  Slot& sl = srv->slots[idx];
  std::unique_lock<std::mutex> lk(srv->mu);
  while (true) {
    // the head of the queue runs first: wait for its arrival time and for
    // room in the budget (a query alone in flight always fits)
    if (srv->queue.empty()) {
      if (srv->stop) return;
      srv->cv_work.wait(lk);
      continue;
    }
    const uint64_t t = srv->queue.front();
    Req& rq = srv->reqs[t];
    const int64_t now = now_ns();
    if (rq.not_before > now) {
      srv->cv_work.wait_until(lk, std::chrono::steady_clock::time_point(
                                      std::chrono::nanoseconds(rq.not_before)));
      continue;
    }
    if (srv->inflight > 0 && srv->inflight + rq.need > srv->budget) {
      srv->cv_work.wait(lk);
      continue;
    }
    srv->queue.pop_front();
    srv->inflight += rq.need;
    const mtl_query q = rq.q;
    const uint64_t need = rq.need;
    lk.unlock();
    srv->cv_work.notify_one();  // the next head may be admissible too
    mtl_query_result r{};
    r.t_start_ns = now_ns();
    int32_t st;
    if (q.kind == MTL_Q_SCAN) {
      st = run_scan(sl, q, r);
    } else {
      uint32_t n = 0;
      st = mtl_star_topk(&q.star, q.star_out, &n, sl.s);
      r.n_rows = n;
    }
    r.status = st;
    if (st != MTL_OK) snprintf(r.error, sizeof(r.error), "%s", last_error());
    r.t_done_ns = now_ns();
    lk.lock();
    srv->inflight -= need;
    Req& done = srv->reqs[t];
    r.t_submit_ns = done.res.t_submit_ns;
    done.res = r;
    done.done = true;
    srv->cv_done.notify_all();
    srv->cv_work.notify_all();
  }
}

}  // namespace mtl

extern "C" {

int64_t mtl_now_ns(void) { return mtl::now_ns(); }

uint64_t mtl_query_scratch_bytes(const mtl_query* q) { return q ? mtl::query_scratch(*q) : 0; }

int32_t mtl_server_create(uint32_t n_slots, uint64_t scratch_budget, mtl_server** out) {
  using namespace mtl;
  if (!out || n_slots == 0 || n_slots > 64)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "n_slots must be 1..64");
  MTL_TRY(ensure_device());
  auto* srv = new mtl_server();
  uint64_t budget = scratch_budget;
  if (!budget) budget = runtime().processing_bytes ? runtime().processing_bytes : (16ull << 30);
  srv->budget = budget;
  srv->slots.resize(n_slots);
  for (auto& sl : srv->slots) {
    int32_t st = mtl_stream_create(&sl.s);
    if (st == MTL_OK && hipMalloc((void**)&sl.words, 64) != hipSuccess)
      st = set_error(MTL_ERR_PROCESSING, "processing error: hipMalloc");
    if (st == MTL_OK && hipHostMalloc(&sl.host, 64, hipHostMallocDefault) != hipSuccess)
      st = set_error(MTL_ERR_PROCESSING, "processing error: hipHostMalloc");
    if (st != MTL_OK) {
      mtl_server_destroy(srv);
      return st;
    }
  }
  for (size_t i = 0; i < srv->slots.size(); ++i) srv->workers.emplace_back(worker, srv, i);
  *out = srv;
  return MTL_OK;
}

int32_t mtl_server_submit(mtl_server* srv, const mtl_query* q, int64_t not_before_ns,
                          uint64_t* ticket) {
  using namespace mtl;
  if (!srv || !q || !ticket) return set_error(MTL_ERR_INVALID_ARGUMENT, "NULL argument");
  if (q->kind != MTL_Q_SCAN && q->kind != MTL_Q_STAR)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "query kind %d", q->kind);
  if (q->kind == MTL_Q_STAR && !q->star_out)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "star query without an output buffer");
  Req rq;
  rq.q = *q;
  rq.need = query_scratch(*q);
  const int64_t now = now_ns();
  rq.not_before = not_before_ns;
  rq.res.t_submit_ns = not_before_ns > now ? not_before_ns : now;
  std::lock_guard<std::mutex> lk(srv->mu);
  const uint64_t t = srv->next_ticket++;
  if (rq.need > srv->budget) {
    rq.done = true;
    rq.res.status = MTL_ERR_BUDGET;
    snprintf(rq.res.error, sizeof(rq.res.error), "query needs more device scratch than the budget");
    rq.res.t_start_ns = rq.res.t_done_ns = now;
    srv->reqs.emplace(t, rq);
  } else {
    srv->reqs.emplace(t, rq);
    srv->queue.push_back(t);
    srv->cv_work.notify_one();
  }
  *ticket = t;
  return MTL_OK;
}

int32_t mtl_server_wait(mtl_server* srv, uint64_t ticket, mtl_query_result* out) {
  using namespace mtl;
  if (!srv || !out) return set_error(MTL_ERR_INVALID_ARGUMENT, "NULL argument");
  std::unique_lock<std::mutex> lk(srv->mu);
  auto it = srv->reqs.find(ticket);
  if (it == srv->reqs.end()) return set_error(MTL_ERR_INVALID_ARGUMENT, "unknown ticket");
  // references into the map survive rehashing by later submits; iterators
  // do not
  mtl::Req& rq = it->second;
  srv->cv_done.wait(lk, [&] { return rq.done; });
  *out = rq.res;
  srv->reqs.erase(ticket);
  if (out->status == MTL_ERR_BUDGET)
    return set_error(MTL_ERR_BUDGET, "query needs more device scratch than the budget");
  return MTL_OK;
}

int32_t mtl_server_destroy(mtl_server* srv) {
  using namespace mtl;
  if (!srv) return MTL_OK;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;
  }
  srv->cv_work.notify_all();
  for (auto& w : srv->workers) w.join();
  for (auto& sl : srv->slots) {
    if (sl.s) hipStreamSynchronize(sl.s->stream);
    if (sl.sel) hipFree(sl.sel);
    if (sl.proj) hipFree(sl.proj);
    if (sl.words) hipFree(sl.words);
    if (sl.host) hipHostFree(sl.host);
    if (sl.s) mtl_stream_destroy(sl.s);
  }
  delete srv;
  return MTL_OK;
}

}  // extern "C"
