// Autotuner: the launch parameters the operators read from the tuned table
// (mtl_internal.hpp TuneId) are timed here, candidate by candidate, on
// synthetic device data of the shapes the serving loop and the SQL lowering
// produce, and the fastest is kept.  Every candidate must give the same
// result as the default (counts, fingerprints, join pairs, group top-K rows
// compared exactly); one that does not is an error, not a choice.  The
// table is saved and loaded as "name=value" lines.
//
// The reference fixes these parameters in code (dense-vs-hash cut-offs and
// load factors in join_plan.rs:29-56, threadgroup sizes in
// aggregate_ops.rs:780-786); this library picks them per device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "device.hpp"
#include "filter.hpp"
#include "mtl_internal.hpp"

namespace mtl {
using namespace dev;

static const TuneParam kTuneDefaults[T_COUNT] = {
    {"small_block", 256, {128, 256, 512, 1024, 0}},
    {"small_steps", 4, {1, 2, 4, 0}},
    {"join_load_pct", 50, {25, 35, 50, 70, 0}},
    {"group_load_pct", 50, {25, 35, 50, 70, 0}},
    {"part_bits", 7, {5, 6, 7, 8, 9, 10, 0}},
};

static TuneParam g_tune[T_COUNT] = {
    kTuneDefaults[0], kTuneDefaults[1], kTuneDefaults[2], kTuneDefaults[3], kTuneDefaults[4],
};

TuneParam* tune_table() { return g_tune; }

static int tune_find(const char* name) {
  for (int i = 0; i < T_COUNT; ++i)
    if (std::strcmp(g_tune[i].name, name) == 0) return i;
  return -1;
}

static bool tune_allowed(int id, int32_t v) {
  for (int c = 0; c < 8 && g_tune[id].candidates[c]; ++c)
    if (g_tune[id].candidates[c] == v) return true;
  return false;
}

static int32_t tune_set(int id, int32_t v) {
  if (!tune_allowed(id, v))
    return set_error(MTL_ERR_INVALID_ARGUMENT, "%s=%d is not a candidate value",
                     g_tune[id].name, v);
  g_tune[id].value = v;
  return MTL_OK;
}

// ---------------------------------------------------------------------------
// Synthetic data
// ---------------------------------------------------------------------------
// mode 0: 64-bit mix of (i + seed); 1: that mod `mod`; 2: a skewed key in
// [0, mod) (the product of two uniform draws, scaled); 3: 32-bit, mix mod
// `mod`; 4: 64-bit, (i mod mod) mixed (mod distinct values).
__global__ void k_tune_fill(void* out, uint64_t n, uint64_t seed, uint64_t mod, int mode) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t x = fp_mix(i + seed);
    uint64_t v;
    switch (mode) {
      case 0: v = x; break;
      case 1: v = x % mod; break;
      case 2: v = ((x & 0xFFFFFFFFull) % mod) * ((x >> 32) % mod) / mod; break;
      case 4: v = fp_mix(i % mod + seed); break;
      default: v = x % mod; break;
    }
    if (mode == 3) ((uint32_t*)out)[i] = (uint32_t)v;
    else ((uint64_t*)out)[i] = v;
  }
}

struct DevBlock {
  void* p = nullptr;
  uint64_t bytes = 0;
  ~DevBlock() { dev_free(p, bytes); }
  int32_t alloc(uint64_t b) {
    bytes = b;
    return dev_alloc(b, &p, "autotune");
  }
};

static int32_t fill(void* out, uint64_t n, uint64_t seed, uint64_t mod, int mode,
                    hipStream_t st) {
  const uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(n, 256), 4096);
  hipLaunchKernelGGL(k_tune_fill, dim3(grid), dim3(256), 0, st, out, n, seed, mod, mode);
  MTL_HIP(hipGetLastError());
  return MTL_OK;
}

static double median(std::vector<double> v) {
  std::sort(v.begin(), v.end());
  return v[v.size() / 2];
}

static double now_us() {
  return std::chrono::duration<double, std::micro>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct Report {
  std::string text;
  void add(const char* group, const char* name, int32_t v, double us, bool chosen) {
    char line[160];
    snprintf(line, sizeof(line), "%s: %s=%d %.1f us%s\n", group, name, v, us,
             chosen ? " (chosen)" : "");
    text += line;
  }
};

// ---------------------------------------------------------------------------
// Small scans: 200 scans of 1,000 to 65,536 rows over an INTEGER and a
// BIGINT column (one or two predicates, projections), one persistent launch;
// candidates are (block, steps) pairs.  The grid is kSmallMax blocks: one
// per query of the largest batch.
// ---------------------------------------------------------------------------
static int32_t tune_small(hipStream_t st, Report& rep) {
  const uint64_t n_max = kSmallRows;
  const uint32_t nq = 200;
  DevBlock dv, db, dblk, dres, dout;
  MTL_TRY(dv.alloc(n_max * 4));
  MTL_TRY(db.alloc(n_max * 8));
  MTL_TRY(fill(dv.p, n_max, 11, 1u << 20, 3, st));
  MTL_TRY(fill(db.p, n_max, 12, 0, 0, st));
  std::vector<mtl_query> qs(nq);
  std::vector<SmallScanItem> items(nq);
  uint64_t out_bytes = 0;
  for (uint32_t i = 0; i < nq; ++i) {
    mtl_query& q = qs[i];
    q = mtl_query{};
    q.kind = MTL_Q_SCAN;
    const double u = (double)(fp_mix(i + 99) % 10000) / 10000.0;
    q.n_rows = std::max<uint64_t>(1000, (uint64_t)(std::pow(10.0, 3.0 + u * 1.8)));
    q.n_rows = std::min<uint64_t>(q.n_rows, n_max);
    const int64_t lit = (int64_t)(fp_mix(i + 7) % (1u << 20));
    q.preds[0] = mtl_pred{dv.p, MTL_INTEGER, MTL_CMP_GT, lit};
    q.n_preds = 1;
    if (i % 3 == 1) {
      q.preds[1] = mtl_pred{db.p, MTL_BIGINT, MTL_CMP_LT, 0};
      q.n_preds = 2;
    }
    if (i % 2 == 0) {
      q.proj_col = dv.p;
      q.proj_type = MTL_INTEGER;
      q.proj_op = MTL_PROJ_MUL;
      q.proj_scalar = 10;
    } else {
      q.proj_col = db.p;
      q.proj_type = MTL_BIGINT;
      q.proj_op = MTL_PROJ_ADD;
      q.proj_scalar = 7;
    }
    out_bytes += q.n_rows * 12;
  }
  MTL_TRY(dout.alloc(out_bytes));
  uint64_t at = 0;
  for (uint32_t i = 0; i < nq; ++i) {
    items[i].q = &qs[i];
    items[i].out_sel = (uint32_t*)((char*)dout.p + at);
    items[i].out_proj = (char*)dout.p + at + qs[i].n_rows * 4;
    at += qs[i].n_rows * 12;
  }
  const size_t bb = small_block_bytes();
  std::vector<char> hblk(bb);
  MTL_TRY(small_scan_stage(hblk.data(), items.data(), nq));
  MTL_TRY(dblk.alloc(bb));
  MTL_TRY(dres.alloc(nq * 16));
  std::vector<uint32_t> ref, got(nq * 4);
  hipEvent_t e0, e1;
  MTL_HIP(hipEventCreate(&e0));
  MTL_HIP(hipEventCreate(&e1));
  struct Ev {
    hipEvent_t a, b;
    ~Ev() {
      hipEventDestroy(a);
      hipEventDestroy(b);
    }
  } ev{e0, e1};
  // one timed launch of (block, steps); results into got
  auto run = [&](int32_t block, int32_t steps, double* us) -> int32_t {
    std::vector<double> t;
    for (int r = 0; r < 4; ++r) {
      MTL_HIP(hipMemcpyAsync(dblk.p, hblk.data(), bb, hipMemcpyHostToDevice, st));
      MTL_HIP(hipEventRecord(e0, st));
      MTL_TRY(small_scan_launch(block, steps, kSmallMax, dblk.p, (uint32_t*)dres.p, st));
      MTL_HIP(hipEventRecord(e1, st));
      MTL_HIP(hipEventSynchronize(e1));
      float ms = 0;
      MTL_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (r) t.push_back(ms * 1e3);
    }
    MTL_HIP(hipMemcpy(got.data(), dres.p, nq * 16, hipMemcpyDeviceToHost));
    *us = median(t);
    return MTL_OK;
  };
  double ref_us = 0;
  MTL_TRY(run(kTuneDefaults[T_SMALL_BLOCK].value, kTuneDefaults[T_SMALL_STEPS].value, &ref_us));
  ref = got;
  int32_t best[2] = {kTuneDefaults[T_SMALL_BLOCK].value, kTuneDefaults[T_SMALL_STEPS].value};
  double best_us = ref_us;
  for (int bi = 0; g_tune[T_SMALL_BLOCK].candidates[bi]; ++bi)
    for (int si = 0; g_tune[T_SMALL_STEPS].candidates[si]; ++si) {
      const int32_t b = g_tune[T_SMALL_BLOCK].candidates[bi];
      const int32_t s = g_tune[T_SMALL_STEPS].candidates[si];
      double us = 0;
      MTL_TRY(run(b, s, &us));
      if (got != ref)
        return set_error(MTL_ERR_PROCESSING,
                         "autotune: small scans with block %d steps %d differ", b, s);
      char nm[64];
      snprintf(nm, sizeof(nm), "block=%d steps", b);
      rep.add("small_scan", nm, s, us, false);
      if (us < best_us) {
        best_us = us;
        best[0] = b;
        best[1] = s;
      }
    }
  MTL_TRY(tune_set(T_SMALL_BLOCK, best[0]));
  MTL_TRY(tune_set(T_SMALL_STEPS, best[1]));
  char nm[64];
  snprintf(nm, sizeof(nm), "block=%d steps", best[0]);
  rep.add("small_scan", nm, best[1], best_us, true);
  return MTL_OK;
}

// ---------------------------------------------------------------------------
// Hash join: 1M unique BIGINT build keys, 8M probe keys (half of them
// match), build + probe per table fill.
// ---------------------------------------------------------------------------
static int32_t tune_join(mtl_stream* s, Report& rep) {
  hipStream_t st = s->stream;
  const uint64_t nb = 1u << 20, np = 8u << 20;
  DevBlock bk, pk, op, ob, oc, dup;
  MTL_TRY(bk.alloc(nb * 8));
  MTL_TRY(pk.alloc(np * 8));
  MTL_TRY(op.alloc(np * 4));
  MTL_TRY(ob.alloc(np * 4));
  MTL_TRY(oc.alloc(16));
  MTL_TRY(dup.alloc(16));
  MTL_TRY(fill(bk.p, nb, 0, nb, 4, st));      // fp_mix(i): unique
  MTL_TRY(fill(pk.p, np, 0, 2 * nb, 4, st));  // fp_mix(i mod 2M): half match
  MTL_HIP(hipStreamSynchronize(st));
  uint64_t ref_n = 0, ref_sum = 0;
  bool have_ref = false;
  int32_t best = kTuneDefaults[T_JOIN_LOAD].value;
  double best_us = 1e300;
  const int32_t saved = g_tune[T_JOIN_LOAD].value;
  for (int ci = 0; g_tune[T_JOIN_LOAD].candidates[ci]; ++ci) {
    const int32_t pct = g_tune[T_JOIN_LOAD].candidates[ci];
    g_tune[T_JOIN_LOAD].value = pct;
    const uint64_t tb = mtl_join_table_bytes(MTL_JOIN_HASH, nb, 0);
    DevBlock table;
    MTL_TRY(table.alloc(tb));
    std::vector<double> t;
    uint32_t cnt = 0;
    for (int r = 0; r < 4; ++r) {
      MTL_HIP(hipMemsetAsync(dup.p, 0, 16, st));
      MTL_HIP(hipStreamSynchronize(st));
      const double t0 = now_us();
      MTL_TRY(mtl_join_build(MTL_JOIN_HASH, bk.p, MTL_BIGINT, nullptr, nullptr, nb, 0, 0,
                             table.p, tb, (uint32_t*)dup.p, s));
      MTL_TRY(mtl_join_probe(MTL_JOIN_HASH, table.p, tb, 0, nb, pk.p, MTL_BIGINT, nullptr,
                             nullptr, np, (uint32_t*)op.p, (uint32_t*)ob.p, (uint32_t*)oc.p, s));
      MTL_HIP(hipStreamSynchronize(st));
      if (r) t.push_back(now_us() - t0);
    }
    MTL_HIP(hipMemcpy(&cnt, oc.p, 4, hipMemcpyDeviceToHost));
    std::vector<uint32_t> hb(cnt);
    if (cnt) MTL_HIP(hipMemcpy(hb.data(), ob.p, (size_t)cnt * 4, hipMemcpyDeviceToHost));
    uint64_t sum = 0;
    for (uint32_t i = 0; i < cnt; ++i) sum = sum * 1000003ull + hb[i];
    if (!have_ref) {
      ref_n = cnt;
      ref_sum = sum;
      have_ref = true;
    } else if (cnt != ref_n || sum != ref_sum) {
      g_tune[T_JOIN_LOAD].value = saved;
      return set_error(MTL_ERR_PROCESSING, "autotune: join at %d%% fill differs", pct);
    }
    const double us = median(t);
    rep.add("hash_join", "join_load_pct", pct, us, false);
    if (us < best_us) {
      best_us = us;
      best = pct;
    }
  }
  MTL_TRY(tune_set(T_JOIN_LOAD, best));
  rep.add("hash_join", "join_load_pct", best, best_us, true);
  return MTL_OK;
}

// ---------------------------------------------------------------------------
// Group-by: the hash table (8M rows, 500K distinct full-range BIGINT keys)
// per fill, and the spill partitioning (16M rows over a skewed 4M-key dense
// domain) per fan-out; top-10 groups by COUNT and the group count compared.
// ---------------------------------------------------------------------------
static int32_t group_run(const mtl_group_desc& d, mtl_stream* s, double* us,
                         std::vector<int64_t>* sig) {
  std::vector<double> t;
  for (int r = 0; r < 3; ++r) {
    MTL_HIP(hipStreamSynchronize(s->stream));
    const double t0 = now_us();
    mtl_groups* g = nullptr;
    MTL_TRY(mtl_group_agg(&d, &g, s));
    uint64_t n = 0;
    int32_t st = mtl_groups_count(g, &n);
    int64_t keys[10] = {0}, cnt[10] = {0}, mx[10] = {0};
    void* outs[2] = {cnt, mx};
    uint32_t got = 0;
    if (st == MTL_OK) st = mtl_groups_topk(g, 0, 1, 10, keys, outs, &got);
    mtl_groups_release(g);
    MTL_TRY(st);
    if (r) t.push_back(now_us() - t0);
    sig->assign({(int64_t)n, (int64_t)got});
    for (uint32_t i = 0; i < got; ++i) {
      sig->push_back(keys[i]);
      sig->push_back(cnt[i]);
      sig->push_back(mx[i]);
    }
  }
  *us = median(t);
  return MTL_OK;
}

static int32_t tune_group(mtl_stream* s, Report& rep) {
  hipStream_t st = s->stream;
  const uint64_t nh = 8u << 20, ns = 16u << 20;
  DevBlock hk, sk, val;
  MTL_TRY(hk.alloc(nh * 8));
  MTL_TRY(sk.alloc(ns * 8));
  MTL_TRY(val.alloc(ns * 8));
  MTL_TRY(fill(hk.p, nh, 5, 500000, 4, st));   // 500K distinct full-range keys
  MTL_TRY(fill(sk.p, ns, 6, 4u << 20, 2, st));  // skewed keys in [0, 4M)
  MTL_TRY(fill(val.p, ns, 7, 1u << 30, 1, st));
  MTL_HIP(hipStreamSynchronize(st));
  auto desc = [&](const void* keys, uint64_t n, int64_t kmin, int64_t kmax) {
    mtl_group_desc d{};
    d.n_rows = n;
    d.n_keys = 1;
    d.key_col[0] = keys;
    d.key_type[0] = MTL_BIGINT;
    d.key_min[0] = kmin;
    d.key_max[0] = kmax;
    d.n_aggs = 2;
    d.aggs[0] = mtl_agg_spec{MTL_AGG_COUNT, MTL_BIGINT, nullptr, 0};
    d.aggs[1] = mtl_agg_spec{MTL_AGG_MAX, MTL_BIGINT, val.p, 0};
    return d;
  };
  // hash table fill
  {
    const mtl_group_desc d = desc(hk.p, nh, INT64_MIN, INT64_MAX);
    const int32_t saved = g_tune[T_GROUP_LOAD].value;
    std::vector<int64_t> ref, sig;
    int32_t best = saved;
    double best_us = 1e300;
    for (int ci = 0; g_tune[T_GROUP_LOAD].candidates[ci]; ++ci) {
      const int32_t pct = g_tune[T_GROUP_LOAD].candidates[ci];
      g_tune[T_GROUP_LOAD].value = pct;
      double us = 0;
      const int32_t rs = group_run(d, s, &us, &sig);
      if (rs != MTL_OK) {
        g_tune[T_GROUP_LOAD].value = saved;
        return rs;
      }
      if (ref.empty()) ref = sig;
      else if (sig != ref) {
        g_tune[T_GROUP_LOAD].value = saved;
        return set_error(MTL_ERR_PROCESSING, "autotune: group-by at %d%% fill differs", pct);
      }
      rep.add("group_hash", "group_load_pct", pct, us, false);
      if (us < best_us) {
        best_us = us;
        best = pct;
      }
    }
    MTL_TRY(tune_set(T_GROUP_LOAD, best));
    rep.add("group_hash", "group_load_pct", best, best_us, true);
  }
  // spill partitioning fan-out
  {
    const mtl_group_desc d = desc(sk.p, ns, 0, (4 << 20) - 1);
    const int32_t saved = g_tune[T_PART_BITS].value;
    std::vector<int64_t> ref, sig;
    int32_t best = saved;
    double best_us = 1e300;
    for (int ci = 0; g_tune[T_PART_BITS].candidates[ci]; ++ci) {
      const int32_t bits = g_tune[T_PART_BITS].candidates[ci];
      g_tune[T_PART_BITS].value = bits;
      double us = 0;
      const int32_t rs = group_run(d, s, &us, &sig);
      if (rs != MTL_OK) {
        g_tune[T_PART_BITS].value = saved;
        return rs;
      }
      if (ref.empty()) ref = sig;
      else if (sig != ref) {
        g_tune[T_PART_BITS].value = saved;
        return set_error(MTL_ERR_PROCESSING, "autotune: partition bits %d differ", bits);
      }
      rep.add("group_spill", "part_bits", bits, us, false);
      if (us < best_us) {
        best_us = us;
        best = bits;
      }
    }
    MTL_TRY(tune_set(T_PART_BITS, best));
    rep.add("group_spill", "part_bits", best, best_us, true);
  }
  return MTL_OK;
}

}  // namespace mtl

using namespace mtl;

extern "C" {

int32_t mtl_tune_get(const char* name, int32_t* value) {
  if (!name || !value) return set_error(MTL_ERR_INVALID_ARGUMENT, "NULL argument");
  const int id = tune_find(name);
  if (id < 0) return set_error(MTL_ERR_INVALID_ARGUMENT, "unknown parameter '%s'", name);
  *value = g_tune[id].value;
  return MTL_OK;
}

int32_t mtl_tune_set(const char* name, int32_t value) {
  if (!name) return set_error(MTL_ERR_INVALID_ARGUMENT, "NULL argument");
  const int id = tune_find(name);
  if (id < 0) return set_error(MTL_ERR_INVALID_ARGUMENT, "unknown parameter '%s'", name);
  return tune_set(id, value);
}

int32_t mtl_tune_reset(void) {
  for (int i = 0; i < T_COUNT; ++i) g_tune[i].value = kTuneDefaults[i].value;
  return MTL_OK;
}

int32_t mtl_tune_describe(char* buf, size_t len) {
  if (!buf || !len) return set_error(MTL_ERR_INVALID_ARGUMENT, "NULL buffer");
  std::string s;
  for (int i = 0; i < T_COUNT; ++i) {
    if (i) s += ' ';
    s += g_tune[i].name;
    s += '=';
    s += std::to_string(g_tune[i].value);
  }
  if (s.size() + 1 > len) return set_error(MTL_ERR_INVALID_ARGUMENT, "buffer too small");
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return MTL_OK;
}

int32_t mtl_tune_save(const char* path) {
  if (!path) return set_error(MTL_ERR_INVALID_ARGUMENT, "NULL path");
  FILE* f = std::fopen(path, "w");
  if (!f) return set_error(MTL_ERR_INVALID_ARGUMENT, "cannot write '%s'", path);
  std::fprintf(f, "# mettle_amd tuned launch parameters (mtl_autotune)\n");
  for (int i = 0; i < T_COUNT; ++i) std::fprintf(f, "%s=%d\n", g_tune[i].name, g_tune[i].value);
  std::fclose(f);
  return MTL_OK;
}

// Every line is checked before any value changes: a bad file leaves the
// table as it was.
int32_t mtl_tune_load(const char* path) {
  if (!path) return set_error(MTL_ERR_INVALID_ARGUMENT, "NULL path");
  FILE* f = std::fopen(path, "r");
  if (!f) return set_error(MTL_ERR_INVALID_ARGUMENT, "cannot read '%s'", path);
  int32_t vals[T_COUNT];
  for (int i = 0; i < T_COUNT; ++i) vals[i] = g_tune[i].value;
  char line[256];
  int ln = 0;
  int32_t st = MTL_OK;
  while (st == MTL_OK && std::fgets(line, sizeof(line), f)) {
    ++ln;
    std::string s(line);
    while (!s.empty() && (s.back() == '\n' || s.back() == '\r' || s.back() == ' ')) s.pop_back();
    size_t b = 0;
    while (b < s.size() && s[b] == ' ') ++b;
    s = s.substr(b);
    if (s.empty() || s[0] == '#') continue;
    const size_t eq = s.find('=');
    if (eq == std::string::npos) {
      st = set_error(MTL_ERR_INVALID_ARGUMENT, "%s:%d: expected name=value", path, ln);
      break;
    }
    const std::string name = s.substr(0, eq), v = s.substr(eq + 1);
    const int id = tune_find(name.c_str());
    char* end = nullptr;
    const long x = std::strtol(v.c_str(), &end, 10);
    if (id < 0) st = set_error(MTL_ERR_INVALID_ARGUMENT, "%s:%d: unknown parameter '%s'", path,
                               ln, name.c_str());
    else if (v.empty() || *end != '\0' || !tune_allowed(id, (int32_t)x))
      st = set_error(MTL_ERR_INVALID_ARGUMENT, "%s:%d: %s=%s is not a candidate value", path,
                     ln, name.c_str(), v.c_str());
    else vals[id] = (int32_t)x;
  }
  std::fclose(f);
  MTL_TRY(st);
  for (int i = 0; i < T_COUNT; ++i) g_tune[i].value = vals[i];
  return MTL_OK;
}

int32_t mtl_autotune(uint32_t groups, char* report, size_t len) {
  MTL_TRY(ensure_device());
  mtl_stream* s = nullptr;
  MTL_TRY(mtl_stream_create(&s));
  struct Close {
    mtl_stream* s;
    ~Close() { mtl_stream_destroy(s); }
  } close{s};
  Report rep;
  int32_t st = MTL_OK;
  if (st == MTL_OK && (groups & MTL_TUNE_SMALL_SCAN)) st = tune_small(s->stream, rep);
  if (st == MTL_OK && (groups & MTL_TUNE_JOIN)) st = tune_join(s, rep);
  if (st == MTL_OK && (groups & MTL_TUNE_GROUP)) st = tune_group(s, rep);
  if (report && len) {
    const size_t n = std::min(len - 1, rep.text.size());
    std::memcpy(report, rep.text.data(), n);
    report[n] = '\0';
  }
  return st;
}

}  // extern "C"
