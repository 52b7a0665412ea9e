// Table ingestion: Arrow record batches -> HBM column cache -> SQL catalog.
//
// Reference: gpu_register_table (src/pinned.rs:187-453): a DuckDB scan is
// read as Arrow batches, each supported column is copied into one host
// vector, VARCHAR is dictionary-encoded, and the table is published under
// its lower-case name; columns of unsupported types or with NULLs are
// skipped (queries naming them fall back to DuckDB's CPU path).
// Differences by design:
//   - DECIMAL stays a scaled int64 (pinned.rs:187-215 casts it to DOUBLE),
//     TIMESTAMP is int64 microseconds (pinned.rs:445-451 skips it);
//   - columns with NULLs are kept: a validity byte per row (1 = valid) is
//     uploaded beside them and the value of a NULL row is written as zero;
//   - the dictionary of a VARCHAR column is sorted, so codes order like the
//     strings (pinned.rs:402-444 numbers values in order of appearance), and
//     it is 1 byte per row when it has at most 256 values;
//   - uploads go through the HBM column cache (mtl_cache_put: pinned
//     staging chunks, the cache budget, reference counts); the catalog
//     keeps the references until the table is dropped or ingested again.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <string_view>
#include <thread>
#include <unordered_map>
#include <vector>

#include "mtl_internal.hpp"

namespace mtl {
namespace {

std::mutex g_ing_mu;
// table -> (cache table id, cached column ids) whose references it holds
std::map<std::string, std::pair<uint64_t, std::vector<uint32_t>>>& ingested() {
  static std::map<std::string, std::pair<uint64_t, std::vector<uint32_t>>> m;
  return m;
}
uint64_t g_version = 0;

std::string lower_name(const char* s) {
  std::string o(s ? s : "");
  for (char& c : o) c = (char)std::tolower((unsigned char)c);
  return o;
}

uint64_t table_key(const std::string& name) {  // FNV-1a
  uint64_t h = 1469598103934665603ull;
  for (unsigned char c : name) h = (h ^ c) * 1099511628211ull;
  return h;
}

bool bit(const void* bitmap, int64_t i) {
  return !bitmap || ((((const uint8_t*)bitmap)[i >> 3] >> (i & 7)) & 1);
}

// One column of the batches, converted for upload.
struct ColOut {
  std::string name;
  int32_t type = 0, scale = 0;
  std::vector<uint8_t> data;
  const void* direct = nullptr;  // one batch needing no conversion: its own value buffer
  uint64_t bytes = 0;            // of the values
  std::vector<uint8_t> valid;  // one byte per row; empty until a NULL is seen
  bool any_null = false;
  int64_t min = INT64_MAX, max = INT64_MIN;
  std::vector<std::string> dict;
  std::string skip;  // non-empty: not ingested, and why
};

enum Src { S_I8, S_I16, S_I32, S_I64, S_U8, S_U16, S_U32, S_U64, S_F32, S_F64, S_DATE32,
           S_DATE64, S_TS, S_DEC, S_UTF8, S_LUTF8, S_NONE };

struct SrcType {
  Src src = S_NONE;
  int64_t ts_mul = 1;  // timestamp unit -> microseconds
  int32_t scale = 0;
};

SrcType parse_format(const char* f, std::string& why) {
  SrcType t;
  const std::string s(f ? f : "");
  static const std::pair<const char*, Src> simple[] = {
      {"c", S_I8}, {"s", S_I16}, {"i", S_I32}, {"l", S_I64}, {"C", S_U8}, {"S", S_U16},
      {"I", S_U32}, {"L", S_U64}, {"f", S_F32}, {"g", S_F64}, {"tdD", S_DATE32},
      {"tdm", S_DATE64}, {"u", S_UTF8}, {"U", S_LUTF8}};
  for (auto& p : simple)
    if (s == p.first) {
      t.src = p.second;
      return t;
    }
  if (s.rfind("ts", 0) == 0 && s.size() >= 4 && s[3] == ':') {
    t.src = S_TS;
    switch (s[2]) {
      case 's': t.ts_mul = 1000000; break;
      case 'm': t.ts_mul = 1000; break;
      case 'u': break;
      default: why = "timestamp unit " + s; t.src = S_NONE;  // ns: DuckDB's TIMESTAMP_NS
    }
    return t;
  }
  if (s.rfind("d:", 0) == 0) {
    int prec = 0, scale = 0, bits = 128;
    const int got = std::sscanf(s.c_str() + 2, "%d,%d,%d", &prec, &scale, &bits);
    if (got >= 2 && bits == 128 && scale >= 0 && scale <= 18) {
      t.src = S_DEC;
      t.scale = scale;
      return t;
    }
    why = "decimal " + s;
    return t;
  }
  why = "Arrow type " + s;
  return t;
}

bool is_index(Src s) { return s == S_I8 || s == S_I16 || s == S_I32 || s == S_I64 ||
                              s == S_U8 || s == S_U16 || s == S_U32; }

int64_t load_int(const void* buf, Src s, int64_t i) {
  switch (s) {
    case S_I8: return ((const int8_t*)buf)[i];
    case S_I16: return ((const int16_t*)buf)[i];
    case S_I32: case S_DATE32: return ((const int32_t*)buf)[i];
    case S_U8: return ((const uint8_t*)buf)[i];
    case S_U16: return ((const uint16_t*)buf)[i];
    case S_U32: return ((const uint32_t*)buf)[i];
    default: return ((const int64_t*)buf)[i];  // I64, U64, DATE64, TS
  }
}

std::string_view utf8_at(const ArrowArray* a, bool large, int64_t i) {
  const char* chars = (const char*)a->buffers[2];
  if (large) {
    const int64_t* o = (const int64_t*)a->buffers[1];
    return std::string_view(chars + o[i], (size_t)(o[i + 1] - o[i]));
  }
  const int32_t* o = (const int32_t*)a->buffers[1];
  return std::string_view(chars + o[i], (size_t)(o[i + 1] - o[i]));
}

// Convert column `ci` of every batch.
void convert(const ArrowSchema* cs, ArrowArray* const* batches, uint32_t nb, int ci,
             uint64_t rows, ColOut& out) {
  out.name = lower_name(cs->name);
  std::string why;
  const bool dict_enc = cs->dictionary != nullptr;
  SrcType st = parse_format(cs->format, why);
  SrcType vt;  // dictionary value type
  if (dict_enc) {
    vt = parse_format(cs->dictionary->format, why);
    if (!is_index(st.src) || (vt.src != S_UTF8 && vt.src != S_LUTF8)) {
      out.skip = "dictionary of " + std::string(cs->dictionary->format ? cs->dictionary->format : "?");
      return;
    }
  } else if (st.src == S_NONE) {
    out.skip = why;
    return;
  }
  const bool strings = dict_enc || st.src == S_UTF8 || st.src == S_LUTF8;
  auto for_rows = [&](auto&& fn) {  // fn(global row, array, index in array, valid)
    uint64_t r0 = 0;
    for (uint32_t b = 0; b < nb; ++b) {
      const ArrowArray* a = batches[b]->children[ci];
      const int64_t base = batches[b]->offset + a->offset;
      const void* vbits = a->null_count == 0 ? nullptr : a->buffers[0];
      for (int64_t i = 0; i < batches[b]->length; ++i)
        fn(r0 + (uint64_t)i, a, base + i, bit(vbits, base + i));
      r0 += (uint64_t)batches[b]->length;
    }
  };
  auto mark_null = [&](uint64_t r) {
    if (!out.any_null) {
      out.valid.assign(rows, 1);
      out.any_null = true;
    }
    out.valid[r] = 0;
  };
  if (strings) {
    // distinct values, sorted (codes order like the strings)
    std::unordered_map<std::string_view, uint32_t> code;
    auto value = [&](const ArrowArray* a, int64_t i) -> std::string_view {
      if (!dict_enc) return utf8_at(a, st.src == S_LUTF8, i);
      const ArrowArray* d = a->dictionary;
      const int64_t k = load_int(a->buffers[1], st.src, i);
      return utf8_at(d, vt.src == S_LUTF8, d->offset + k);
    };
    for_rows([&](uint64_t, const ArrowArray* a, int64_t i, bool ok) {
      if (ok) code.emplace(value(a, i), 0);
    });
    std::vector<std::string_view> vals;
    vals.reserve(code.size());
    for (auto& kv : code) vals.push_back(kv.first);
    std::sort(vals.begin(), vals.end());
    for (uint32_t k = 0; k < vals.size(); ++k) code[vals[k]] = k;
    out.dict.assign(vals.begin(), vals.end());
    const bool narrow = vals.size() <= 256;
    out.type = narrow ? MTL_DICT8 : MTL_DICT32;
    out.data.assign(rows * (narrow ? 1 : 4), 0);
    out.bytes = out.data.size();
    for_rows([&](uint64_t r, const ArrowArray* a, int64_t i, bool ok) {
      if (!ok) return mark_null(r);
      const uint32_t c = code[value(a, i)];
      if (narrow) out.data[r] = (uint8_t)c;
      else std::memcpy(out.data.data() + r * 4, &c, 4);
    });
    out.min = 0;
    out.max = vals.empty() ? 0 : (int64_t)vals.size() - 1;
    return;
  }
  // a batch without NULLs whose values keep their width: one copy of its
  // value range (and, for integers, a min/max loop the compiler vectorises)
  auto bulk = [&](int w, bool minmax) {
    uint64_t r0 = 0;
    for (uint32_t b = 0; b < nb; ++b) {
      const ArrowArray* a = batches[b]->children[ci];
      if (a->null_count != 0 && a->buffers[0]) return false;
    }
    out.bytes = rows * (uint64_t)w;
    if (nb == 1) {  // upload straight from the batch's buffer
      const ArrowArray* a = batches[0]->children[ci];
      out.direct = (const uint8_t*)a->buffers[1] + (batches[0]->offset + a->offset) * w;
      out.data.clear();
      out.data.shrink_to_fit();
    } else {
      out.data.resize(out.bytes);
    }
    for (uint32_t b = 0; b < nb; ++b) {
      const ArrowArray* a = batches[b]->children[ci];
      const int64_t base = batches[b]->offset + a->offset, len = batches[b]->length;
      if (!out.direct)
        std::memcpy(out.data.data() + r0 * w, (const uint8_t*)a->buffers[1] + base * w,
                    (size_t)len * w);
      if (minmax) {
        int64_t lo = out.min, hi = out.max;
        if (w == 4) {
          const int32_t* v = (const int32_t*)((const uint8_t*)a->buffers[1]) + base;
          for (int64_t i = 0; i < len; ++i) {
            lo = v[i] < lo ? v[i] : lo;
            hi = v[i] > hi ? v[i] : hi;
          }
        } else {
          const int64_t* v = (const int64_t*)((const uint8_t*)a->buffers[1]) + base;
          for (int64_t i = 0; i < len; ++i) {
            lo = v[i] < lo ? v[i] : lo;
            hi = v[i] > hi ? v[i] : hi;
          }
        }
        out.min = lo;
        out.max = hi;
      }
      r0 += (uint64_t)len;
    }
    return true;
  };
  if (st.src == S_F32 || st.src == S_F64) {
    const int w = st.src == S_F32 ? 4 : 8;
    out.type = st.src == S_F32 ? MTL_FLOAT : MTL_DOUBLE;
    if (bulk(w, false)) {
      out.min = out.max = 0;
      return;
    }
    out.data.assign(rows * w, 0);
    out.bytes = rows * (uint64_t)w;
    for_rows([&](uint64_t r, const ArrowArray* a, int64_t i, bool ok) {
      if (!ok) return mark_null(r);
      std::memcpy(out.data.data() + r * w, (const uint8_t*)a->buffers[1] + i * w, w);
    });
    out.min = out.max = 0;
    return;
  }
  if (st.src == S_DEC) {
    out.type = MTL_DECIMAL;
    out.scale = st.scale;
    out.data.assign(rows * 8, 0);
    out.bytes = out.data.size();
    for_rows([&](uint64_t r, const ArrowArray* a, int64_t i, bool ok) {
      if (!out.skip.empty()) return;
      if (!ok) return mark_null(r);
      int64_t lo, hi;
      std::memcpy(&lo, (const uint8_t*)a->buffers[1] + i * 16, 8);
      std::memcpy(&hi, (const uint8_t*)a->buffers[1] + i * 16 + 8, 8);
      if (hi != (lo < 0 ? -1 : 0)) {
        out.skip = "decimal value beyond int64";
        return;
      }
      std::memcpy(out.data.data() + r * 8, &lo, 8);
      out.min = std::min(out.min, lo);
      out.max = std::max(out.max, lo);
    });
    return;
  }
  // integers, dates, timestamps
  bool wide = st.src == S_I64 || st.src == S_U32 || st.src == S_U64 || st.src == S_TS ||
              st.src == S_DATE64;
  out.type = st.src == S_DATE32 || st.src == S_DATE64 ? MTL_DATE
             : st.src == S_TS                         ? MTL_TIMESTAMP
             : wide                                   ? MTL_BIGINT
                                                      : MTL_INTEGER;
  if (st.src == S_DATE64) wide = false;
  const int w = wide ? 8 : 4;
  const bool same = st.src == S_I32 || st.src == S_DATE32 || st.src == S_I64 ||
                    (st.src == S_TS && st.ts_mul == 1);
  if (same && bulk(w, true)) return;
  out.data.assign(rows * w, 0);
  out.bytes = rows * (uint64_t)w;
  out.min = INT64_MAX;
  out.max = INT64_MIN;
  for_rows([&](uint64_t r, const ArrowArray* a, int64_t i, bool ok) {
    if (!out.skip.empty()) return;
    if (!ok) return mark_null(r);
    int64_t v = load_int(a->buffers[1], st.src, i);
    if (st.src == S_U64 && v < 0) {
      out.skip = "uint64 value beyond int64";
      return;
    }
    if (st.src == S_DATE64) {  // milliseconds -> days, rounding down
      v = v >= 0 ? v / 86400000 : -((-v + 86399999) / 86400000);
    } else if (st.src == S_TS && st.ts_mul != 1) {
      if (__builtin_mul_overflow(v, st.ts_mul, &v)) {
        out.skip = "timestamp beyond int64 microseconds";
        return;
      }
    }
    if (w == 4) {
      const int32_t x = (int32_t)v;
      std::memcpy(out.data.data() + r * 4, &x, 4);
    } else {
      std::memcpy(out.data.data() + r * 8, &v, 8);
    }
    out.min = std::min(out.min, v);
    out.max = std::max(out.max, v);
  });
}

void release_refs(const std::string& table) {
  auto it = ingested().find(table);
  if (it == ingested().end()) return;
  for (uint32_t c : it->second.second) mtl_cache_release(it->second.first, c);
  ingested().erase(it);
}

}  // namespace

void ingest_release(const char* table) {
  std::lock_guard<std::mutex> lk(g_ing_mu);
  if (table) {
    release_refs(lower_name(table));
  } else {
    while (!ingested().empty()) release_refs(ingested().begin()->first);
  }
}

}  // namespace mtl

extern "C" int32_t mtl_catalog_ingest_arrow(const char* table, const ArrowSchema* schema,
                                            ArrowArray* const* batches, uint32_t n_batches,
                                            const char* const* unique_cols, uint32_t n_unique,
                                            uint64_t* out_rows, uint64_t* out_bytes,
                                            char* skipped, size_t skipped_len) {
  using namespace mtl;
  if (!table || !schema || (n_batches && !batches))
    return set_error(MTL_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!schema->format || std::strcmp(schema->format, "+s") != 0)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "schema is not a struct (record batch)");
  uint64_t rows = 0;
  for (uint32_t b = 0; b < n_batches; ++b) {
    if (!batches[b] || batches[b]->n_children != schema->n_children)
      return set_error(MTL_ERR_INVALID_ARGUMENT, "batch %u does not match the schema", b);
    rows += (uint64_t)batches[b]->length;
  }
  if (rows > 0xFFFFFFFFull) return set_error(MTL_UNSUPPORTED, "more than 2^32-1 rows");
  const std::string name = lower_name(table);
  std::lock_guard<std::mutex> lk(g_ing_mu);
  release_refs(name);
  const uint64_t tid = table_key(name);
  const uint64_t version = ++g_version;
  MTL_TRY(mtl_catalog_add_table(name.c_str(), rows));
  std::vector<uint32_t> held;
  auto fail = [&](int32_t st) {
    for (uint32_t c : held) mtl_cache_release(tid, c);
    catalog_drop_table(name.c_str());
    return st;
  };
  auto upload = [&](uint32_t col_id, const void* host, uint64_t bytes,
                    const void** dev) -> int32_t {
    static const uint8_t zeros[16] = {0};
    if (!bytes) host = zeros;
    MTL_TRY(mtl_cache_put(tid, col_id, host, bytes ? bytes : 16, version, dev));
    held.push_back(col_id);
    return MTL_OK;
  };
  uint64_t total = 0;
  std::string skip;
  // convert the columns on host threads (at most 8 at a time), then upload
  const int64_t nc = schema->n_children;
  std::vector<ColOut> conv((size_t)nc);
  for (int64_t c0 = 0; c0 < nc; c0 += 8) {
    std::vector<std::thread> th;
    for (int64_t ci = c0; ci < std::min<int64_t>(nc, c0 + 8); ++ci)
      th.emplace_back([&, ci] {
        convert(schema->children[ci], batches, n_batches, (int)ci, rows, conv[(size_t)ci]);
      });
    for (auto& t : th) t.join();
  }
  for (int64_t ci = 0; ci < nc; ++ci) {
    ColOut c = std::move(conv[(size_t)ci]);
    if (!c.skip.empty()) {
      skip += (skip.empty() ? "" : ",") + c.name;
      continue;
    }
    if (c.min > c.max) c.min = c.max = 0;  // no valid value
    const void* dev = nullptr;
    const void* vdev = nullptr;
    int32_t st = upload((uint32_t)ci, c.direct ? c.direct : (const void*)c.data.data(), c.bytes,
                        &dev);
    if (st == MTL_OK && c.any_null)
      st = upload((uint32_t)ci | 0x80000000u, c.valid.data(), c.valid.size(), &vdev);
    bool unique = false;
    for (uint32_t u = 0; u < n_unique; ++u)
      if (unique_cols && unique_cols[u] && lower_name(unique_cols[u]) == c.name) unique = true;
    if (st == MTL_OK)
      st = mtl_catalog_add_column(name.c_str(), c.name.c_str(), c.type, c.scale, dev, c.min,
                                  c.max, unique ? 1 : 0);
    if (st == MTL_OK && c.any_null)
      st = mtl_catalog_set_validity(name.c_str(), c.name.c_str(), (const uint8_t*)vdev);
    if (st == MTL_OK && !c.dict.empty()) {
      std::vector<const char*> vals;
      for (auto& v : c.dict) vals.push_back(v.c_str());
      st = mtl_catalog_set_dictionary(name.c_str(), c.name.c_str(), vals.data(),
                                      (uint32_t)vals.size());
    }
    if (st != MTL_OK) return fail(st);
    total += c.bytes + c.valid.size();
  }
  ingested()[name] = {tid, held};
  if (out_rows) *out_rows = rows;
  if (out_bytes) *out_bytes = total;
  if (skipped && skipped_len) std::snprintf(skipped, skipped_len, "%s", skip.c_str());
  return MTL_OK;
}
