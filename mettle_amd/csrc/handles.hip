// Refcounted device handles: mtl_buf (a device allocation) and mtl_vec (a
// selection vector: uint32 row ids, a device-resident count and a host
// capacity).
//
// Reference: GpuBuffer (gpu_buffer.rs:12-18), a refcounted Metal buffer
// with len and element size whose Clone shares the buffer, and GpuVector
// (gpu_buffer.rs:46-52), which adds the device count and a host max_len;
// the count is read only after end_batch (gpu_buffer.rs:94-103).  Here the
// handles own memory charged to the processing budget (freed at the last
// release), the operators take their device pointers, and mtl_vec_count is
// the one synchronising read of a count.
#include <atomic>

#include "mtl_internal.hpp"

struct mtl_buf {
  std::atomic<int64_t> refs{1};
  void* dev = nullptr;
  uint64_t bytes = 0;
};

struct mtl_vec {
  std::atomic<int64_t> refs{1};
  uint32_t* rows = nullptr;   // capacity entries
  uint32_t* count = nullptr;  // one device word (the next 255 bytes: padding)
  uint64_t capacity = 0;
  void* mem = nullptr;        // one allocation: count word, then the rows
  uint64_t bytes = 0;
};

extern "C" {

int32_t mtl_buf_alloc(uint64_t bytes, mtl_buf** out) {
  using namespace mtl;
  if (!out) return set_error(MTL_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  MTL_TRY(ensure_device());
  mtl_buf* b = new mtl_buf();
  b->bytes = bytes;
  const int32_t st = dev_alloc(bytes ? bytes : 1, &b->dev, "mtl_buf");
  if (st != MTL_OK) {
    delete b;
    return st;
  }
  *out = b;
  return MTL_OK;
}

int32_t mtl_buf_retain(mtl_buf* b) {
  if (!b) return mtl::set_error(MTL_ERR_INVALID_ARGUMENT, "buffer is NULL");
  b->refs.fetch_add(1, std::memory_order_relaxed);
  return MTL_OK;
}

int32_t mtl_buf_release(mtl_buf* b) {
  if (!b) return MTL_OK;
  if (b->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) {
    mtl::dev_free(b->dev, b->bytes ? b->bytes : 1);
    delete b;
  }
  return MTL_OK;
}

void* mtl_buf_device(const mtl_buf* b) { return b ? b->dev : nullptr; }
uint64_t mtl_buf_bytes(const mtl_buf* b) { return b ? b->bytes : 0; }

int32_t mtl_vec_alloc(uint64_t capacity, mtl_vec** out) {
  using namespace mtl;
  if (!out) return set_error(MTL_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  if (capacity > 0xFFFFFFFFull) return set_error(MTL_UNSUPPORTED, "more than 2^32-1 rows");
  MTL_TRY(ensure_device());
  mtl_vec* v = new mtl_vec();
  v->capacity = capacity;
  v->bytes = 256 + std::max<uint64_t>(capacity, 1) * 4;
  int32_t st = dev_alloc(v->bytes, &v->mem, "mtl_vec");
  if (st == MTL_OK) st = hipMemset(v->mem, 0, 4) == hipSuccess
                             ? MTL_OK
                             : set_error(MTL_ERR_PROCESSING, "mtl_vec: clearing the count failed");
  if (st != MTL_OK) {
    if (v->mem) dev_free(v->mem, v->bytes);
    delete v;
    return st;
  }
  v->count = (uint32_t*)v->mem;
  v->rows = (uint32_t*)((char*)v->mem + 256);
  *out = v;
  return MTL_OK;
}

int32_t mtl_vec_retain(mtl_vec* v) {
  if (!v) return mtl::set_error(MTL_ERR_INVALID_ARGUMENT, "vector is NULL");
  v->refs.fetch_add(1, std::memory_order_relaxed);
  return MTL_OK;
}

int32_t mtl_vec_release(mtl_vec* v) {
  if (!v) return MTL_OK;
  if (v->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) {
    mtl::dev_free(v->mem, v->bytes);
    delete v;
  }
  return MTL_OK;
}

uint32_t* mtl_vec_rows(const mtl_vec* v) { return v ? v->rows : nullptr; }
uint32_t* mtl_vec_count_ptr(const mtl_vec* v) { return v ? v->count : nullptr; }
uint64_t mtl_vec_capacity(const mtl_vec* v) { return v ? v->capacity : 0; }

int32_t mtl_vec_count(const mtl_vec* v, mtl_stream* s, uint64_t* n) {
  using namespace mtl;
  if (!v || !n) return set_error(MTL_ERR_INVALID_ARGUMENT, "NULL argument");
  uint32_t c = 0;
  MTL_TRY(mtl_memcpy_d2h(&c, v->count, 4, s));
  if (c > v->capacity)
    return set_error(MTL_ERR_PROCESSING, "selection count %u exceeds its capacity %llu", c,
                     (unsigned long long)v->capacity);
  *n = c;
  return MTL_OK;
}

int32_t mtl_filter_vec(const mtl_pred* preds, uint32_t n_preds, uint64_t n_rows,
                       const mtl_vec* in, mtl_vec* out, mtl_stream* s) {
  using namespace mtl;
  if (!out) return set_error(MTL_ERR_INVALID_ARGUMENT, "out is NULL");
  const uint64_t cap = in ? in->capacity : n_rows;
  if (out->capacity < cap)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "output vector holds %llu rows, needs %llu",
                     (unsigned long long)out->capacity, (unsigned long long)cap);
  return mtl_filter(preds, n_preds, n_rows, in ? in->rows : nullptr, in ? in->count : nullptr,
                    in ? in->capacity : 0, out->rows, out->count, s);
}

}  // extern "C"
