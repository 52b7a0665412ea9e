// Runtime: errors, budgets (gpu_buffer_init), device bring-up, stream
// handles and their scratch arenas, device memory helpers.
//
// Reference counterparts: src/config.rs (budget parsing), src/state.rs
// (initialise / not-initialised error), src/metal/runtime/mod.rs:72-151
// (device + queue creation), 182-236 (batching -> streams here).
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>

#include "mtl_internal.hpp"
#include "prof.hpp"

namespace mtl {

namespace {
thread_local std::string g_last_error;
}

int32_t set_error(int32_t code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

const char* last_error() { return g_last_error.c_str(); }

Runtime& runtime() {
  static Runtime* rt = new Runtime();  // never destroyed: outlives atexit
  return *rt;
}

int32_t charge_processing(uint64_t bytes, const char* what) {
  Runtime& rt = runtime();
  const uint64_t limit = rt.initialized ? rt.processing_bytes : 0;
  uint64_t cur = rt.device_bytes_in_use.load();
  while (true) {
    if (limit && (cur + bytes < cur || cur + bytes > limit))
      return set_error(MTL_ERR_BUDGET,
                       "processing budget exceeded: %s needs %llu bytes with %llu of %llu "
                       "in use (gpu_buffer_init)",
                       what, (unsigned long long)bytes, (unsigned long long)cur,
                       (unsigned long long)limit);
    if (rt.device_bytes_in_use.compare_exchange_weak(cur, cur + bytes)) return MTL_OK;
  }
}

void release_processing(uint64_t bytes) { runtime().device_bytes_in_use -= bytes; }

int32_t dev_alloc(uint64_t bytes, void** out, const char* what) {
  MTL_TRY(charge_processing(bytes, what));
  hipError_t e = hipMalloc(out, bytes ? bytes : 1);
  if (e != hipSuccess) {
    release_processing(bytes);
    *out = nullptr;
    return set_error(MTL_ERR_PROCESSING, "processing error: %s hipMalloc(%llu): %s", what,
                     (unsigned long long)bytes, hipGetErrorString(e));
  }
  return MTL_OK;
}

void dev_free(void* p, uint64_t bytes) {
  if (!p) return;
  hipFree(p);
  release_processing(bytes);
}

int32_t Arena::alloc(size_t bytes, void** out) {
  size_t need = (bytes + 255) & ~size_t(255);
  if (need == 0) need = 256;
  if (chunks.empty() || off + need > chunks.back().cap) {
    // doubling chunks, or exactly what is needed when the doubled chunk
    // would not fit the processing budget
    size_t cap = chunks.empty() ? (size_t(64) << 20) : chunks.back().cap * 2;
    while (cap < need) cap *= 2;
    if (charge_processing(cap, "scratch arena") != MTL_OK) {
      cap = (need + (size_t(2) << 20) - 1) & ~((size_t(2) << 20) - 1);
      MTL_TRY(charge_processing(cap, "scratch arena"));
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, cap);
    if (e != hipSuccess) {
      release_processing(cap);
      return set_error(MTL_ERR_PROCESSING,
                       "processing error: scratch arena hipMalloc(%zu): %s",
                       cap, hipGetErrorString(e));
    }
    chunks.push_back({(char*)p, cap});
    off = 0;
  }
  *out = chunks.back().base + off;
  off += need;
  if (off > high_water) high_water = off;
  return MTL_OK;
}

void Arena::release_retired() {
  while (chunks.size() > 1) {
    dev_free(chunks.front().base, chunks.front().cap);
    chunks.erase(chunks.begin());
  }
}

void Arena::release_all() {
  for (auto& c : chunks) dev_free(c.base, c.cap);
  chunks.clear();
  off = 0;
}

int32_t ensure_device() {
  Runtime& rt = runtime();
  std::lock_guard<std::mutex> g(rt.mu);
  if (rt.default_stream) return MTL_OK;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0)
    return set_error(MTL_ERR_DEVICE_UNAVAILABLE,
                     "metal unavailable: no HIP device is visible (%s)",
                     e == hipSuccess ? "0 devices" : hipGetErrorString(e));
  MTL_HIP(hipSetDevice(rt.device));
  hipDeviceProp_t prop;
  MTL_HIP(hipGetDeviceProperties(&prop, rt.device));
  rt.device_name = prop.name;
  rt.arch_name = prop.gcnArchName;
  rt.cu_count = prop.multiProcessorCount;
  // stream-ordered allocations (group tables) stay in the device's pool when
  // freed instead of being unmapped at the next synchronisation, so a table
  // of hundreds of MB costs no new mapping per query; the processing budget
  // still accounts for every live byte
  hipMemPool_t pool;
  if (hipDeviceGetDefaultMemPool(&pool, rt.device) == hipSuccess) {
    uint64_t keep = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  }
  mtl_stream* s = new mtl_stream();
  e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete s;
    return set_error(MTL_ERR_DEVICE_UNAVAILABLE,
                     "metal unavailable: hipStreamCreate: %s",
                     hipGetErrorString(e));
  }
  s->owns_stream = true;
  rt.default_stream = s;
  return MTL_OK;
}

int32_t resolve_stream(mtl_stream* s, mtl_stream** out) {
  if (s) {
    *out = s;
    return MTL_OK;
  }
  MTL_TRY(ensure_device());
  *out = runtime().default_stream;
  return MTL_OK;
}

int32_t stream_pinned(mtl_stream* s, size_t bytes, void** out) {
  if (s->pinned_small_bytes < bytes) {
    if (s->pinned_small) hipHostFree(s->pinned_small);
    s->pinned_small = nullptr;
    size_t cap = 4096;
    while (cap < bytes) cap *= 2;
    MTL_HIP(hipHostMalloc(&s->pinned_small, cap, hipHostMallocDefault));
    s->pinned_small_bytes = cap;
  }
  *out = s->pinned_small;
  return MTL_OK;
}

// parse_human_bytes semantics (src/config.rs:61-123): digits and '.' form
// the number, whitespace is ignored, the rest is a 1024-based unit.
static int32_t parse_bytes(const char* raw, uint64_t* out) {
  if (!raw) return set_error(MTL_ERR_INVALID_BUFFER_CONFIG,
                             "invalid buffer config: buffer sizes cannot be empty");
  std::string s(raw);
  size_t a = s.find_first_not_of(" \t\r\n");
  size_t b = s.find_last_not_of(" \t\r\n");
  if (a == std::string::npos)
    return set_error(MTL_ERR_INVALID_BUFFER_CONFIG,
                     "invalid buffer config: buffer sizes cannot be empty");
  s = s.substr(a, b - a + 1);
  std::string num, unit;
  for (char ch : s) {
    if ((ch >= '0' && ch <= '9') || ch == '.') {
      if (!unit.empty())
        return set_error(MTL_ERR_INVALID_BUFFER_CONFIG,
                         "invalid buffer config: invalid size literal '%s'", raw);
      num.push_back(ch);
    } else if (ch == ' ' || ch == '\t' || ch == '\r' || ch == '\n') {
      continue;
    } else {
      unit.push_back(ch);
    }
  }
  if (num.empty())
    return set_error(MTL_ERR_INVALID_BUFFER_CONFIG,
                     "invalid buffer config: missing numeric value in '%s'", raw);
  // exactly one optional '.', digits elsewhere (Rust f64::from_str accepts
  // "1.", ".5", rejects "1.2.3")
  size_t dots = 0;
  for (char ch : num) dots += ch == '.';
  if (dots > 1 || num == ".")
    return set_error(MTL_ERR_INVALID_BUFFER_CONFIG,
                     "invalid buffer config: invalid numeric value in '%s'", raw);
  double q = strtod(num.c_str(), nullptr);
  if (!std::isfinite(q) || q < 0)
    return set_error(MTL_ERR_INVALID_BUFFER_CONFIG,
                     "invalid buffer config: invalid size literal '%s'", raw);
  std::string u;
  for (char ch : unit) u.push_back((char)tolower((unsigned char)ch));
  double mult;
  if (u.empty() || u == "b") mult = 1;
  else if (u == "k" || u == "kb" || u == "kib") mult = 1024.0;
  else if (u == "m" || u == "mb" || u == "mib") mult = 1024.0 * 1024.0;
  else if (u == "g" || u == "gb" || u == "gib") mult = 1024.0 * 1024.0 * 1024.0;
  else if (u == "t" || u == "tb" || u == "tib")
    mult = 1024.0 * 1024.0 * 1024.0 * 1024.0;
  else
    return set_error(MTL_ERR_INVALID_BUFFER_CONFIG,
                     "invalid buffer config: unsupported size unit '%s' in '%s'",
                     u.c_str(), raw);
  double bytes = q * mult;
  if (bytes > 18446744073709551615.0)
    return set_error(MTL_ERR_INVALID_BUFFER_CONFIG,
                     "invalid buffer config: size literal '%s' overflowed u64", raw);
  *out = (uint64_t)llround(bytes);
  return MTL_OK;
}

}  // namespace mtl

using namespace mtl;

extern "C" {

int32_t mtl_parse_bytes(const char* text, uint64_t* out_bytes) {
  if (!out_bytes) return set_error(MTL_ERR_INVALID_ARGUMENT, "out_bytes is NULL");
  return parse_bytes(text, out_bytes);
}

int32_t mtl_init(uint64_t cache_bytes, uint64_t processing_bytes,
                 uint64_t pinned_bytes, char* device_name,
                 size_t device_name_len) {
  if (pinned_bytes == 0) {
    if (cache_bytes + processing_bytes < cache_bytes)
      return set_error(MTL_ERR_INVALID_BUFFER_CONFIG,
                       "invalid buffer config: pinned memory default overflowed "
                       "the configured budgets");
    pinned_bytes = cache_bytes + processing_bytes;
  }
  if (cache_bytes == 0)
    return set_error(MTL_ERR_INVALID_BUFFER_CONFIG,
                     "invalid buffer config: gpu_buffer_init requires a non-zero "
                     "cache region");
  if (processing_bytes == 0)
    return set_error(MTL_ERR_INVALID_BUFFER_CONFIG,
                     "invalid buffer config: gpu_buffer_init requires a non-zero "
                     "processing region");
  MTL_TRY(ensure_device());
  Runtime& rt = runtime();
  {
    std::lock_guard<std::mutex> g(rt.mu);
    rt.cache_bytes = cache_bytes;
    rt.processing_bytes = processing_bytes;
    rt.pinned_bytes = pinned_bytes;
    rt.initialized = true;
  }
  if (device_name && device_name_len) {
    snprintf(device_name, device_name_len, "%s", rt.device_name.c_str());
  }
  return MTL_OK;
}

int32_t mtl_buffer_init(const char* cache, const char* processing,
                        const char* pinned, char* device_name,
                        size_t device_name_len) {
  uint64_t c = 0, p = 0, pin = 0;
  MTL_TRY(parse_bytes(cache, &c));
  MTL_TRY(parse_bytes(processing, &p));
  if (pinned) {
    MTL_TRY(parse_bytes(pinned, &pin));
    if (pin == 0)
      return set_error(MTL_ERR_INVALID_BUFFER_CONFIG,
                       "invalid buffer config: gpu_buffer_init requires a non-zero "
                       "pinned region");
  }
  return mtl_init(c, p, pin, device_name, device_name_len);
}

int32_t mtl_budgets(uint64_t* cache_bytes, uint64_t* processing_bytes,
                    uint64_t* pinned_bytes) {
  Runtime& rt = runtime();
  std::lock_guard<std::mutex> g(rt.mu);
  if (!rt.initialized)
    return set_error(MTL_ERR_NOT_INITIALIZED,
                     "not initialized: call gpu_buffer_init before gpu_processing");
  if (cache_bytes) *cache_bytes = rt.cache_bytes;
  if (processing_bytes) *processing_bytes = rt.processing_bytes;
  if (pinned_bytes) *pinned_bytes = rt.pinned_bytes;
  return MTL_OK;
}

int32_t mtl_shutdown(void) {
  MTL_TRY(mtl_cache_reset());
  Runtime& rt = runtime();
  std::lock_guard<std::mutex> g(rt.mu);
  if (rt.default_stream) {
    mtl_stream* d = rt.default_stream;
    hipStreamSynchronize(d->stream);
    d->arena.release_all();
    for (int i = 0; i < 2; ++i) {
      if (d->parked_table[i]) hipFree(d->parked_table[i]);
      d->parked_table[i] = nullptr;
      d->parked_bytes[i] = 0;
    }
  }
  rt.initialized = false;
  rt.cache_bytes = rt.processing_bytes = rt.pinned_bytes = 0;
  return MTL_OK;
}

int32_t mtl_last_error(char* buf, size_t len) {
  if (buf && len) snprintf(buf, len, "%s", last_error());
  return MTL_OK;
}

int32_t mtl_device_count(int32_t* out) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  if (out) *out = n;
  return MTL_OK;
}

int32_t mtl_stream_create(mtl_stream** out) {
  if (!out) return set_error(MTL_ERR_INVALID_ARGUMENT, "out is NULL");
  MTL_TRY(ensure_device());
  mtl_stream* s = new mtl_stream();
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete s;
    return set_error(MTL_ERR_PROCESSING, "processing error: hipStreamCreate: %s",
                     hipGetErrorString(e));
  }
  s->owns_stream = true;
  *out = s;
  return MTL_OK;
}

int32_t mtl_stream_destroy(mtl_stream* s) {
  if (!s) return MTL_OK;
  if (s == runtime().default_stream)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "cannot destroy the default stream");
  hipStreamSynchronize(s->stream);
  s->arena.release_all();
  for (void* t : s->parked_table)
    if (t) hipFree(t);
  if (s->pinned_small) hipHostFree(s->pinned_small);
  if (s->owns_stream) hipStreamDestroy(s->stream);
  delete s;
  return MTL_OK;
}

int32_t mtl_stream_sync(mtl_stream* s) {
  MTL_TRY(resolve_stream(s, &s));
  MTL_HIP(stream_wait(s->stream));
  prof_collect(s);
  s->arena.release_retired();
  return MTL_OK;
}

void* mtl_stream_native(mtl_stream* s) {
  if (resolve_stream(s, &s) != MTL_OK) return nullptr;
  return (void*)s->stream;
}

namespace {
std::mutex g_alloc_mu;
std::unordered_map<void*, uint64_t> g_allocs;  // mtl_alloc pointer -> bytes charged
}  // namespace

int32_t mtl_alloc(uint64_t bytes, void** out_dev) {
  if (!out_dev) return set_error(MTL_ERR_INVALID_ARGUMENT, "out_dev is NULL");
  MTL_TRY(ensure_device());
  MTL_TRY(dev_alloc(bytes, out_dev, "mtl_alloc"));
  std::lock_guard<std::mutex> g(g_alloc_mu);
  g_allocs[*out_dev] = bytes;
  return MTL_OK;
}

int32_t mtl_free(void* dev) {
  if (!dev) return MTL_OK;
  uint64_t bytes = 0;
  {
    std::lock_guard<std::mutex> g(g_alloc_mu);
    auto it = g_allocs.find(dev);
    if (it == g_allocs.end())
      return set_error(MTL_ERR_INVALID_ARGUMENT, "mtl_free of a pointer mtl_alloc did not return");
    bytes = it->second;
    g_allocs.erase(it);
  }
  dev_free(dev, bytes);
  return MTL_OK;
}

int32_t mtl_processing_in_use(uint64_t* bytes) {
  if (!bytes) return set_error(MTL_ERR_INVALID_ARGUMENT, "bytes is NULL");
  *bytes = runtime().device_bytes_in_use.load();
  return MTL_OK;
}

int32_t mtl_memcpy_h2d(void* dst_dev, const void* src_host, uint64_t bytes,
                       mtl_stream* s) {
  MTL_TRY(resolve_stream(s, &s));
  MTL_HIP(hipMemcpyAsync(dst_dev, src_host, bytes, hipMemcpyHostToDevice,
                         s->stream));
  MTL_HIP(stream_wait(s->stream));
  return MTL_OK;
}

int32_t mtl_memcpy_d2h(void* dst_host, const void* src_dev, uint64_t bytes,
                       mtl_stream* s) {
  MTL_TRY(resolve_stream(s, &s));
  MTL_HIP(hipMemcpyAsync(dst_host, src_dev, bytes, hipMemcpyDeviceToHost,
                         s->stream));
  MTL_HIP(stream_wait(s->stream));
  return MTL_OK;
}

}  // extern "C"
