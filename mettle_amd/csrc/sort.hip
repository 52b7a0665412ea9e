// Device ORDER BY: stable sort of a selection by one key column.
//
// Replaces the CPU comparator sort of the output batch, sort_record_batch
// (helpers.rs:627-667), for the ORDER BY of the scan path.  The keys are
// gathered through the selection and mapped to order-preserving unsigned
// integers (signed integers: sign bit flipped; floats: DuckDB's total order
// with NaN largest and -0 == 0; DESC: complemented), then sorted as (key,
// row id) pairs by rocPRIM's LSD radix sort, which is stable: equal keys
// keep the selection's (ascending row) order, so the result is exactly
// ORDER BY key [DESC], row.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "device.hpp"
#include "mtl_internal.hpp"
#include "prof.hpp"

namespace mtl {
using namespace dev;

template <typename K>
__global__ void k_sort_keys(const void* col, int32_t type, const uint32_t* sel,
                            const uint32_t* count, uint64_t cap, int32_t desc, K* keys,
                            uint32_t* rows) {
  const uint64_t n = dev_count(count, cap);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = sel ? sel[i] : (uint32_t)i;
    K e;
    if constexpr (sizeof(K) == 4)  // 4-byte integer types: sign bit flipped
      e = (uint32_t)(int32_t)load_key(col, type, r) ^ 0x80000000u;
    else
      e = enc_value(col, type, r);
    keys[i] = desc ? (K)~e : e;
    rows[i] = r;
  }
}

static bool narrow_sort_type(int32_t t) {
  return t == MTL_INTEGER || t == MTL_DATE || t == MTL_DICT32 || t == MTL_DICT8;
}

template <typename K>
static int32_t sort_pairs(const void* col, int32_t type, const uint32_t* sel,
                          const uint32_t* count, uint64_t n, int32_t desc, uint32_t* out_rows,
                          mtl_stream* s) {
  hipStream_t st = s->stream;
  Arena& ar = s->arena;
  void *kin, *kout, *rin;
  MTL_TRY(ar.alloc(n * sizeof(K), &kin));
  MTL_TRY(ar.alloc(n * sizeof(K), &kout));
  MTL_TRY(ar.alloc(n * 4, &rin));
  const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, 256 * 8), 4096));
  {
    ProfScope ps(s, "k_sort_keys");
    hipLaunchKernelGGL(k_sort_keys<K>, dim3(grid), dim3(256), 0, st, col, type, sel, count, n,
                       desc, (K*)kin, (uint32_t*)rin);
  }
  MTL_HIP(hipGetLastError());
  size_t temp_bytes = 0;
  MTL_HIP(rocprim::radix_sort_pairs(nullptr, temp_bytes, (K*)kin, (K*)kout, (uint32_t*)rin,
                                    out_rows, (unsigned int)n, 0, (unsigned int)(8 * sizeof(K)),
                                    st));
  void* temp;
  MTL_TRY(ar.alloc(temp_bytes ? temp_bytes : 1, &temp));
  ProfScope ps(s, "radix_sort_pairs");
  MTL_HIP(rocprim::radix_sort_pairs(temp, temp_bytes, (K*)kin, (K*)kout, (uint32_t*)rin,
                                    out_rows, (unsigned int)n, 0, (unsigned int)(8 * sizeof(K)),
                                    st));
  return MTL_OK;
}

}  // namespace mtl

extern "C" int32_t mtl_sort(const void* column, int32_t type, const uint32_t* sel,
                            const uint32_t* count, uint64_t capacity, int32_t desc,
                            uint32_t* out_rows, uint32_t* out_n, mtl_stream* s) {
  using namespace mtl;
  MTL_TRY(resolve_stream(s, &s));
  if (!column || !out_rows) return set_error(MTL_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!type_width(type) || type == MTL_HUGEINT)
    return set_error(MTL_UNSUPPORTED, "ORDER BY on type %d", type);
  if (capacity > 0xFFFFFFFFull) return set_error(MTL_UNSUPPORTED, "more than 2^32-1 rows");
  std::lock_guard<std::mutex> lk(s->mu);
  s->arena.reset();
  uint64_t n = capacity;
  if (sel && count) {  // the row count lives on the device: read it once
    void* host;
    MTL_TRY(stream_pinned(s, 8, &host));
    MTL_HIP(hipMemcpyAsync(host, count, 4, hipMemcpyDeviceToHost, s->stream));
    MTL_HIP(stream_wait(s->stream));
    n = *(const uint32_t*)host;
    if (n > capacity) return set_error(MTL_ERR_INVALID_ARGUMENT, "count exceeds capacity");
  }
  if (out_n) *out_n = (uint32_t)n;
  if (n == 0) return MTL_OK;
  if (narrow_sort_type(type))
    return sort_pairs<uint32_t>(column, type, sel, nullptr, n, desc, out_rows, s);
  return sort_pairs<uint64_t>(column, type, sel, nullptr, n, desc, out_rows, s);
}
