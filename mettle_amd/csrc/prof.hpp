// Kernel-level timing with hipEvents on the stream the kernel runs on.
// Counterpart of the reference's METTLE_PROFILE phase timers
// (src/plan/executor/cache.rs:10-21), which used host clocks only.
// Disabled by default; when enabled every instrumented launch records an
// event pair, and the pairs are resolved at the next stream synchronisation.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

struct mtl_stream;

namespace mtl {

bool prof_enabled();
// Record the start/end event of a named launch on stream s.
void prof_begin(mtl_stream* s, const char* name);
void prof_end(mtl_stream* s);
// Resolve pending pairs of s (call after hipStreamSynchronize).
void prof_collect(mtl_stream* s);

struct ProfScope {
  mtl_stream* s;
  bool on;
  // allow = false: never record (a path that must not create events)
  ProfScope(mtl_stream* s_, const char* name, bool allow = true)
      : s(s_), on(allow && prof_enabled()) {
    if (on) prof_begin(s, name);
  }
  ~ProfScope() {
    if (on) prof_end(s);
  }
};

}  // namespace mtl
