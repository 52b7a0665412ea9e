// Host-visible argument blocks of the filter/compaction kernels, shared by
// the operator entry points and the fused star pipeline.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mettle_hip.h"

namespace mtl {

struct ZeroList {
  uint32_t* ptr[8];
  uint64_t words[8];
  uint32_t value[8];
  int n = 0;
  void add(uint32_t* p, uint64_t w, uint32_t v = 0) {
    ptr[n] = p;
    words[n] = w;
    value[n] = v;
    ++n;
  }
};
int32_t launch_zero(const ZeroList& zl, hipStream_t st);

struct FilterArgs {
  mtl_pred preds[MTL_MAX_PREDS];
  uint32_t n_preds;
  uint32_t vec_ok;  // bit p: column p is aligned for vector loads
  uint64_t n_rows;  // rows (no in_sel) or capacity of in_sel
  const uint32_t* in_sel;
  const uint32_t* in_count;
  uint32_t* out_sel;
  uint32_t* out_count;
  uint64_t* status;   // look-back status, one word per tile (zeroed)
  uint32_t* tile_ctr; // ticket counter (zeroed)
  uint32_t* err;      // bit 0: look-back spin cap, bit 1: duplicate mark key
  // optional semi-join probe: keep the row only if bit (key - min) is set
  const uint64_t* probe_bm;
  const void* probe_key;
  int32_t probe_key_type;
  int64_t probe_min;
  uint64_t probe_span;
  // optional side effect: set bit (key - min) for every kept row
  uint64_t* mark_bm;
  const void* mark_key;
  int32_t mark_key_type;
  int64_t mark_min;
  uint64_t mark_span;
};

uint32_t filter_tiles(uint64_t n);
int32_t launch_filter(FilterArgs a, uint64_t capacity, hipStream_t st);

}  // namespace mtl
