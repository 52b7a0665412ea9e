// hipEvent-based kernel timing (see prof.hpp).
#include <map>

#include "mtl_internal.hpp"
#include "prof.hpp"

namespace mtl {

namespace {
std::atomic<bool> g_prof_on{false};
std::mutex g_prof_mu;
struct Acc {
  double ms = 0;
  uint64_t count = 0;
};
std::map<std::string, Acc>& table() {
  static auto* t = new std::map<std::string, Acc>();
  return *t;
}

hipEvent_t take_event(mtl_stream* s) {
  if (!s->prof_pool.empty()) {
    hipEvent_t e = s->prof_pool.back();
    s->prof_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
}  // namespace

bool prof_enabled() { return g_prof_on.load(std::memory_order_relaxed); }

void prof_begin(mtl_stream* s, const char* name) {
  hipEvent_t a = take_event(s), b = take_event(s);
  if (!a || !b) return;
  hipEventRecord(a, s->stream);
  s->prof_pending.push_back({name, a, b});
}

void prof_end(mtl_stream* s) {
  if (s->prof_pending.empty()) return;
  hipEventRecord(s->prof_pending.back().end, s->stream);
}

void prof_collect(mtl_stream* s) {
  if (s->prof_pending.empty()) return;
  std::lock_guard<std::mutex> g(g_prof_mu);
  for (auto& p : s->prof_pending) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, p.start, p.end) == hipSuccess) {
      Acc& a = table()[p.name];
      a.ms += ms;
      a.count += 1;
    }
    s->prof_pool.push_back(p.start);
    s->prof_pool.push_back(p.end);
  }
  s->prof_pending.clear();
}

}  // namespace mtl

using namespace mtl;

extern "C" {

int32_t mtl_profile_enable(int32_t on) {
  g_prof_on.store(on != 0);
  return MTL_OK;
}

int32_t mtl_profile_reset(void) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  table().clear();
  return MTL_OK;
}

int32_t mtl_profile_get(uint32_t index, char* name, size_t name_len, double* total_ms,
                        uint64_t* count) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  if (index >= table().size())
    return set_error(MTL_ERR_INVALID_ARGUMENT, "profile index %u out of range", index);
  auto it = table().begin();
  std::advance(it, index);
  if (name && name_len) snprintf(name, name_len, "%s", it->first.c_str());
  if (total_ms) *total_ms = it->second.ms;
  if (count) *count = it->second.count;
  return MTL_OK;
}

}  // extern "C"
