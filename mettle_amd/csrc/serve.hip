// Serving loop: queries arrive (optionally at scheduled times), wait in
// arrival order, and are admitted while their device scratch fits the
// budget.  n_slots workers, each with its own stream and arena, take the
// admitted queries in batches: every query that has arrived, up to
// max_batch, joins the batch a free worker forms (continuous batching: a
// query arriving while others run joins the next batch as soon as a worker
// is free, with no batch barrier across workers).  Within a batch,
//   - scans with one predicate on the same column (projecting the same
//     column or nothing) share their two table passes (filter.hpp);
//   - star queries over the same three tables share their dimension,
//     parent and fact passes (star.hpp);
// every shared query's result lands in one device region copied back with
// one host synchronisation for the batch.  Other queries (several
// predicates, group-by, SQL) run one after another on the worker's stream.
// With n_slots == 1 and max_batch == 1 every query runs alone on the
// single-query operators: the reference's serial dispatch
// (materialize_query, processing.rs:103-143, one query per CALL inside
// bind).
//
// Scan queries leave their rows on the device; the host gets the row count
// and an order-sensitive fingerprint of the (row id, value) pairs, computed
// on the device, so the result of a query can be compared alone and in a
// mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <string>
#include <thread>
#include <unordered_map>

#include "device.hpp"
#include "filter.hpp"
#include "mtl_internal.hpp"
#include "star.hpp"

namespace mtl {
using namespace dev;

uint64_t group_scratch_bytes(const mtl_group_desc& d);
int32_t sql_scratch_bytes(const char* sql, uint64_t* bytes);

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
  switch (q.kind) {
    case MTL_Q_SCAN: {
      const uint64_t tiles = ceil_div(q.n_rows ? q.n_rows : 1, kTileRows);
      const uint64_t pw = q.proj_col ? (uint64_t)std::max(type_width(q.proj_type), 4) : 0;
      const uint64_t groups = ceil_div(tiles, 1u << kGroupShift);
      // selection + projection, pass bits and tile counts, group sums and
      // fingerprint sums (a cache line each per group), clearing words
      return q.n_rows * (4 + pw) + tiles * (kTileRows / 8 + 4) + groups * 512 + (1u << 16);
    }
    case MTL_Q_STAR: {
      const mtl_star_desc& d = q.star;
      const uint64_t pspan = key_off(d.p_key_max, d.p_key_min) + 1;
      const uint64_t dspan = d.d_rows ? key_off(d.d_key_max, d.d_key_min) + 1 : 0;
      const uint64_t ftiles = ceil_div(d.f_rows ? d.f_rows : 1, kTileRows);
      return d.p_rows * 20 + ftiles * kTileRows * 8 + pspan / 8 * 2 + dspan / 4 + d.d_rows / 4 +
             (uint64_t)d.k * 1024 * 64 + (1u << 20);
    }
    case MTL_Q_GROUP:
      return group_scratch_bytes(q.group);
    default:
      return 0;  // MTL_Q_SQL: estimated at submit
  }
}

struct Req {
  mtl_query q;
  std::string sql;
  uint64_t need = 0;
  int64_t not_before = 0;
  bool done = false;
  mtl_query_result res{};
};

struct Slot {
  mtl_stream* s = nullptr;
  // single-query scan outputs (grown, charged to the processing budget)
  uint32_t* sel = nullptr;
  void* proj = nullptr;
  uint64_t cap = 0;
  uint32_t* words = nullptr;  // [0] count, [1] overflow flag, [2..3] fingerprint
  void* host = nullptr;       // pinned copy of a batch's results
  size_t host_cap = 0;
  void* stage = nullptr;      // pinned per-query arguments of the shared passes
  size_t stage_cap = 0;
  SmallGraph* small = nullptr;  // persistent small-scan launch (batching servers)
};

static int32_t slot_reserve(Slot& sl, uint64_t rows) {
  if (sl.sel && rows <= sl.cap) return MTL_OK;
  MTL_HIP(hipStreamSynchronize(sl.s->stream));
  dev_free(sl.sel, sl.cap * 4);
  dev_free(sl.proj, sl.cap * 8);
  sl.sel = nullptr;
  sl.proj = nullptr;
  sl.cap = 0;
  const uint64_t cap = std::max<uint64_t>(rows, 1 << 16);
  MTL_TRY(dev_alloc(cap * 4, (void**)&sl.sel, "serving slot"));
  if (int32_t st = dev_alloc(cap * 8, &sl.proj, "serving slot")) {
    dev_free(sl.sel, cap * 4);
    sl.sel = nullptr;
    return st;
  }
  sl.cap = cap;
  return MTL_OK;
}

static int32_t pinned_reserve(void*& p, size_t& cap, size_t bytes) {
  if (p && bytes <= cap) return MTL_OK;
  if (p) hipHostFree(p);
  p = nullptr;
  cap = 0;
  size_t c = 4096;
  while (c < bytes) c *= 2;
  MTL_HIP(hipHostMalloc(&p, c, hipHostMallocDefault));
  cap = c;
  return MTL_OK;
}

// One clearing launch (the filter's, with the slot's words added), the
// filter pass, the expansion with the projection and the fingerprint fused,
// the fingerprint sum, and one 16-byte copy of (count, overflow, fingerprint).
static int32_t run_scan(Slot& sl, const mtl_query& q, mtl_query_result& r) {
  MTL_TRY(slot_reserve(sl, q.n_rows));
  MTL_TRY(pinned_reserve(sl.host, sl.host_cap, 64));
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

static int32_t run_group(Slot& sl, const mtl_query& q, mtl_query_result& r) {
  mtl_groups* g = nullptr;
  MTL_TRY(mtl_group_agg(&q.group, &g, sl.s));
  struct Release {
    mtl_groups* g;
    ~Release() { mtl_groups_release(g); }
  } rel{g};
  MTL_TRY(mtl_groups_count(g, &r.groups));
  if (q.group_k > 0) {
    uint32_t n = 0;
    MTL_TRY(mtl_groups_topk(g, q.group_agg, q.group_desc, q.group_k, q.group_keys,
                            q.group_out, &n));
    r.n_rows = n;
    return MTL_OK;
  }
  if (r.groups > q.group_capacity)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "%llu groups exceed the capacity %llu",
                     (unsigned long long)r.groups, (unsigned long long)q.group_capacity);
  uint64_t n = 0;
  MTL_TRY(mtl_groups_fetch(g, q.group_keys, q.group_out, q.group_capacity, &n));
  r.n_rows = (uint32_t)n;
  return MTL_OK;
}

// One query on the single-query operators.
static int32_t run_one(Slot& sl, const mtl_query& q, mtl_query_result& r) {
  switch (q.kind) {
    case MTL_Q_SCAN:
      return run_scan(sl, q, r);
    case MTL_Q_STAR: {
      uint32_t n = 0;
      int32_t st = mtl_star_topk(&q.star, q.star_out, &n, sl.s);
      r.n_rows = n;
      return st;
    }
    case MTL_Q_GROUP:
      return run_group(sl, q, r);
    default: {
      mtl_result* res = nullptr;
      int32_t st = mtl_sql_execute(q.sql, &res, sl.s);
      r.result = res;
      return st;
    }
  }
}

static void finish(mtl_query_result& r, int32_t st) {
  r.status = st;
  if (st != MTL_OK) snprintf(r.error, sizeof(r.error), "%s", last_error());
}

// A batch of two or more queries: the share groups go through the shared
// passes with one synchronisation; the rest run one by one afterwards.
static void run_batch(Slot& sl, std::vector<mtl_query>& qs, std::vector<mtl_query_result>& rs) {
  const uint32_t n = (uint32_t)qs.size();
  std::vector<int> group(n, -1);  // share group of each shared query
  std::vector<std::vector<uint32_t>> groups;
  std::vector<int> kind;  // per group: MTL_Q_SCAN or MTL_Q_STAR
  // per star group: one query of each distinct dimension predicate
  std::vector<std::vector<uint32_t>> dims;
  // small scans (group -2): one persistent launch from the slot's graph
  std::vector<uint32_t> small;
  for (uint32_t i = 0; i < n; ++i) {
    const mtl_query& q = qs[i];
    if (sl.small && small.size() < kSmallMax && small_scan_ok(q)) {
      group[i] = -2;
      small.push_back(i);
      continue;
    }
    const bool scan = q.kind == MTL_Q_SCAN && scan_batchable(q);
    const bool star = q.kind == MTL_Q_STAR && star_batchable(q.star) && q.star_out;
    if (!scan && !star) continue;
    for (size_t g = 0; g < groups.size() && group[i] < 0; ++g) {
      const mtl_query& h = qs[groups[g][0]];
      const uint32_t cap = scan ? kMQMax : kStarMQMax;
      if (kind[g] != q.kind || groups[g].size() >= cap) continue;
      if (scan) {
        if (!scan_shareable(h, q)) continue;
      } else {
        if (!star_shareable(h.star, q.star)) continue;
        bool known = false;
        for (uint32_t r : dims[g]) known = known || star_dim_same(qs[r].star, q.star);
        if (!known) {
          if (dims[g].size() >= kDimSlots) continue;
          dims[g].push_back(i);
        }
      }
      group[i] = (int)g;
      groups[g].push_back(i);
    }
    if (group[i] < 0) {
      group[i] = (int)groups.size();
      groups.push_back({i});
      kind.push_back(q.kind);
      dims.push_back(star ? std::vector<uint32_t>{i} : std::vector<uint32_t>{});
    }
  }
  if (!groups.empty() || !small.empty()) {
    int32_t st = MTL_OK;
    // device results: scans 16 B each, stars out_n (8 B) + K rows each
    // (assigned group by group: a scan group's words are contiguous)
    std::vector<uint64_t> roff(n, 0);
    uint64_t rbytes = 0, stage_bytes = 0;
    for (const auto& mem : groups)
      for (uint32_t i : mem) {
        roff[i] = rbytes;
        rbytes += qs[i].kind == MTL_Q_SCAN
                      ? 16
                      : (16 + (uint64_t)qs[i].star.k * sizeof(mtl_star_row) + 15) & ~uint64_t(15);
      }
    auto stage_of = [&](size_t g) -> size_t {
      const size_t b = kind[g] == MTL_Q_SCAN ? scan_batch_stage_bytes(groups[g].size())
                                             : star_batch_stage_bytes(groups[g].size());
      return (b + 255) & ~size_t(255);
    };
    for (size_t g = 0; g < groups.size(); ++g) stage_bytes += stage_of(g);
    {
      std::lock_guard<std::mutex> lk(sl.s->mu);
      Arena& ar = sl.s->arena;
      ar.reset();
      void* dres = nullptr;
      st = pinned_reserve(sl.host, sl.host_cap, rbytes);
      if (st == MTL_OK) st = pinned_reserve(sl.stage, sl.stage_cap, stage_bytes);
      if (st == MTL_OK) st = ar.alloc(rbytes, &dres);
      if (st == MTL_OK && !small.empty()) {
        std::vector<SmallScanItem> items(small.size());
        for (size_t j = 0; j < small.size() && st == MTL_OK; ++j) {
          const mtl_query& q = qs[small[j]];
          items[j].q = &q;
          void* o = nullptr;
          st = ar.alloc(q.n_rows * 4, &o);
          items[j].out_sel = (uint32_t*)o;
          items[j].out_proj = nullptr;
          if (st == MTL_OK && q.proj_col) {
            st = ar.alloc(q.n_rows * std::max(type_width(q.proj_type), 4), &o);
            items[j].out_proj = o;
          }
        }
        if (st == MTL_OK) st = small_scan_enqueue(sl.small, items.data(), (uint32_t)items.size(), sl.s);
      }
      char* stage = (char*)sl.stage;
      for (size_t g = 0; g < groups.size() && st == MTL_OK; ++g) {
        const auto& mem = groups[g];
        if (kind[g] == MTL_Q_SCAN) {
          std::vector<ScanBatchItem> items(mem.size());
          for (size_t j = 0; j < mem.size() && st == MTL_OK; ++j) {
            const mtl_query& q = qs[mem[j]];
            items[j].q = &q;
            void* o = nullptr;
            st = ar.alloc(q.n_rows * 4, &o);
            items[j].out_sel = (uint32_t*)o;
            items[j].out_proj = nullptr;
            if (st == MTL_OK && q.proj_col) {
              st = ar.alloc(q.n_rows * std::max(type_width(q.proj_type), 4), &o);
              items[j].out_proj = o;
            }
          }
          // the group's words are contiguous: its members' result slots
          // follow one another in roff order
          if (st == MTL_OK)
            st = scan_batch_enqueue(items.data(), (uint32_t)items.size(),
                                    (uint32_t*)((char*)dres + roff[mem[0]]), stage, sl.s);
          stage += stage_of(g);
        } else {
          std::vector<StarBatchItem> items(mem.size());
          for (size_t j = 0; j < mem.size(); ++j) {
            char* at = (char*)dres + roff[mem[j]];
            items[j].d = &qs[mem[j]].star;
            items[j].out_n = (uint32_t*)at;
            items[j].out = (mtl_star_row*)(at + 16);
          }
          st = star_batch_enqueue(items.data(), (uint32_t)items.size(), stage, sl.s);
          stage += stage_of(g);
        }
      }
      if (st == MTL_OK && rbytes)
        st = hipMemcpyAsync(sl.host, dres, rbytes, hipMemcpyDeviceToHost, sl.s->stream) ==
                     hipSuccess
                 ? MTL_OK
                 : set_error(MTL_ERR_PROCESSING, "processing error: result copy");
      // whatever was enqueued completes before the arena or staging is reused
      const hipError_t se = hipStreamSynchronize(sl.s->stream);
      if (st == MTL_OK && se != hipSuccess)
        st = set_error(MTL_ERR_PROCESSING, "processing error: %s", hipGetErrorString(se));
      ar.release_retired();
    }
    for (size_t j = 0; j < small.size(); ++j) {
      const uint32_t i = small[j];
      mtl_query_result& r = rs[i];
      r.batch = n;
      if (st != MTL_OK) {
        finish(r, st);
        continue;
      }
      const uint32_t* w = small_scan_results(sl.small) + 4 * j;
      r.count = w[0];
      r.fingerprint = (uint64_t)w[2] | ((uint64_t)w[3] << 32);
      int32_t qs_ = MTL_OK;
      if (w[0] > qs[i].n_rows) qs_ = set_error(MTL_ERR_PROCESSING, "device scan failed");
      else if (w[1]) qs_ = set_error(MTL_ERR_OVERFLOW, "Overflow in projection");
      finish(r, qs_);
    }
    for (uint32_t i = 0; i < n; ++i) {
      if (group[i] < 0) continue;
      mtl_query_result& r = rs[i];
      r.batch = n;
      if (st != MTL_OK) {
        finish(r, st);
        continue;
      }
      const char* at = (const char*)sl.host + roff[i];
      const mtl_query& q = qs[i];
      if (q.kind == MTL_Q_SCAN) {
        const uint32_t* w = (const uint32_t*)at;
        r.count = w[0];
        r.fingerprint = (uint64_t)w[2] | ((uint64_t)w[3] << 32);
        int32_t qs_ = MTL_OK;
        if (w[0] > q.n_rows) qs_ = set_error(MTL_ERR_PROCESSING, "device scan failed");
        else if (w[1]) qs_ = set_error(MTL_ERR_OVERFLOW, "Overflow in projection");
        finish(r, qs_);
      } else {
        const uint32_t* w = (const uint32_t*)at;
        int32_t qs_ = star_check_error(w[1]);
        if (qs_ == MTL_OK && w[0] > q.star.k)
          qs_ = set_error(MTL_ERR_PROCESSING, "device star run failed");
        if (qs_ == MTL_OK) {
          r.n_rows = w[0];
          std::memcpy(q.star_out, at + 16, (size_t)w[0] * sizeof(mtl_star_row));
        }
        finish(r, qs_);
      }
    }
  }
  for (uint32_t i = 0; i < n; ++i) {
    if (group[i] != -1) continue;
    rs[i].batch = n;
    finish(rs[i], run_one(sl, qs[i], rs[i]));
  }
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
  uint32_t max_batch = 1;
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
    // the batch: the head, then every later query that has arrived, in
    // arrival order, while the budget holds
    std::vector<uint64_t> tickets{t};
    uint64_t need = rq.need;
    srv->queue.pop_front();
    for (auto it = srv->queue.begin();
         it != srv->queue.end() && tickets.size() < srv->max_batch;) {
      Req& o = srv->reqs[*it];
      if (o.not_before > now || srv->inflight + need + o.need > srv->budget) break;
      need += o.need;
      tickets.push_back(*it);
      it = srv->queue.erase(it);
    }
    srv->inflight += need;
    std::vector<mtl_query> qs;
    for (uint64_t tk : tickets) {
      Req& o = srv->reqs[tk];
      qs.push_back(o.q);
      if (o.q.kind == MTL_Q_SQL) qs.back().sql = o.sql.c_str();  // stable until done
    }
    lk.unlock();
    srv->cv_work.notify_one();  // the next head may be admissible too
    std::vector<mtl_query_result> rs(qs.size());
    const int64_t t_start = now_ns();
    if (qs.size() == 1) {
      rs[0].batch = 1;
      finish(rs[0], run_one(sl, qs[0], rs[0]));
    } else {
      run_batch(sl, qs, rs);
    }
    const int64_t t_done = now_ns();
    lk.lock();
    srv->inflight -= need;
    for (size_t i = 0; i < tickets.size(); ++i) {
      Req& done = srv->reqs[tickets[i]];
      rs[i].t_submit_ns = done.res.t_submit_ns;
      rs[i].t_start_ns = t_start;
      rs[i].t_done_ns = t_done;
      done.res = rs[i];
      done.done = true;
    }
    srv->cv_done.notify_all();
    srv->cv_work.notify_all();
  }
}

static void fail_at_submit(Req& rq, int32_t status, int64_t now) {
  rq.done = true;
  rq.res.status = status;
  snprintf(rq.res.error, sizeof(rq.res.error), "%s", last_error());
  rq.res.t_start_ns = rq.res.t_done_ns = now;
}

static uint64_t enqueue(mtl_server* srv, Req&& rq) {
  std::lock_guard<std::mutex> lk(srv->mu);
  const uint64_t t = srv->next_ticket++;
  auto& slot = srv->reqs.emplace(t, std::move(rq)).first->second;
  if (!slot.done) {
    if (slot.need > srv->budget) {
      set_error(MTL_ERR_BUDGET, "query needs %llu bytes of device scratch, more than the "
                "budget of %llu", (unsigned long long)slot.need,
                (unsigned long long)srv->budget);
      fail_at_submit(slot, MTL_ERR_BUDGET, slot.res.t_submit_ns);
    } else {
      srv->queue.push_back(t);
      srv->cv_work.notify_one();
    }
  }
  return t;
}

}  // namespace mtl

extern "C" {

int64_t mtl_now_ns(void) { return mtl::now_ns(); }

uint64_t mtl_query_scratch_bytes(const mtl_query* q) { return q ? mtl::query_scratch(*q) : 0; }

int32_t mtl_server_create(uint32_t n_slots, uint32_t max_batch, uint64_t scratch_budget,
                          mtl_server** out) {
  using namespace mtl;
  if (!out || n_slots == 0 || n_slots > 64)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "n_slots must be 1..64");
  if (max_batch == 0 || max_batch > 256)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "max_batch must be 1..256");
  MTL_TRY(ensure_device());
  auto* srv = new mtl_server();
  uint64_t budget = scratch_budget;
  if (!budget) budget = runtime().processing_bytes ? runtime().processing_bytes : (16ull << 30);
  srv->budget = budget;
  srv->max_batch = max_batch;
  srv->slots.resize(n_slots);
  for (auto& sl : srv->slots) {
    int32_t st = mtl_stream_create(&sl.s);
    if (st == MTL_OK) st = dev_alloc(64, (void**)&sl.words, "serving slot");
    if (st == MTL_OK) st = pinned_reserve(sl.host, sl.host_cap, 4096);
    if (st == MTL_OK && max_batch > 1) st = small_graph_create(sl.s, &sl.small);
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
  if (q->kind < MTL_Q_SCAN || q->kind > MTL_Q_GROUP)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "query kind %d (SQL: mtl_server_submit_sql)",
                     q->kind);
  if (q->kind == MTL_Q_STAR && !q->star_out)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "star query without an output buffer");
  if (q->kind == MTL_Q_GROUP && !q->group_keys)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "group query without output buffers");
  Req rq;
  rq.q = *q;
  rq.q.sql = nullptr;
  rq.need = query_scratch(*q);
  const int64_t now = now_ns();
  rq.not_before = not_before_ns;
  rq.res.t_submit_ns = not_before_ns > now ? not_before_ns : now;
  *ticket = enqueue(srv, std::move(rq));
  return MTL_OK;
}

int32_t mtl_server_submit_sql(mtl_server* srv, const char* sql, int64_t not_before_ns,
                              uint64_t* ticket) {
  using namespace mtl;
  if (!srv || !sql || !ticket) return set_error(MTL_ERR_INVALID_ARGUMENT, "NULL argument");
  Req rq;
  rq.q = mtl_query{};
  rq.q.kind = MTL_Q_SQL;
  rq.sql = sql;
  const int64_t now = now_ns();
  rq.not_before = not_before_ns;
  rq.res.t_submit_ns = not_before_ns > now ? not_before_ns : now;
  // lowered now: a statement outside the GPU subset completes at once with
  // MTL_UNSUPPORTED (DuckDB's CPU path), and the plan sizes the admission
  const int32_t st = sql_scratch_bytes(sql, &rq.need);
  if (st != MTL_OK) fail_at_submit(rq, st, now);
  *ticket = enqueue(srv, std::move(rq));
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
  if (out->status == MTL_ERR_BUDGET) return set_error(MTL_ERR_BUDGET, "%s", out->error);
  return MTL_OK;
}

int32_t mtl_server_poll(mtl_server* srv, uint64_t ticket, mtl_query_result* out,
                        int32_t* done) {
  using namespace mtl;
  if (!srv || !out || !done) return set_error(MTL_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(srv->mu);
  auto it = srv->reqs.find(ticket);
  if (it == srv->reqs.end()) return set_error(MTL_ERR_INVALID_ARGUMENT, "unknown ticket");
  *done = it->second.done ? 1 : 0;
  if (!*done) return MTL_OK;
  *out = it->second.res;
  srv->reqs.erase(it);
  if (out->status == MTL_ERR_BUDGET) return set_error(MTL_ERR_BUDGET, "%s", out->error);
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
    dev_free(sl.sel, sl.cap * 4);
    dev_free(sl.proj, sl.cap * 8);
    dev_free(sl.words, 64);
    if (sl.host) hipHostFree(sl.host);
    if (sl.stage) hipHostFree(sl.stage);
    small_graph_destroy(sl.small);
    if (sl.s) mtl_stream_destroy(sl.s);
  }
  // results nobody waited for: their SQL result handles
  for (auto& kv : srv->reqs)
    if (kv.second.res.result) mtl_result_free(kv.second.res.result);
  delete srv;
  return MTL_OK;
}

}  // extern "C"
