/*
 * mettle_oracle.c — CPU restatement of the mettle hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file computes expected results for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only to compute expected results or a CPU timing.  The
 * product path (libmettle_hip.so and mettle_amd/) never calls it.
 *
 * Semantics follow the reference where the reference matches DuckDB, and
 * DuckDB where it does not (the reference computed in float32):
 *   - predicates: kernels.metal:70-173 (compare_*_scalar) with DuckDB's float
 *     total order (NaN equal to NaN and larger than every other value);
 *   - compaction: row ids of passing rows in ascending order (the closed form
 *     of filter_to_indices, kernels.metal:684-696, minus its atomic order);
 *   - projection: kernels.metal:4-68 op codes with exact integer arithmetic
 *     and DuckDB's overflow error;
 *   - joins: dense lookup (kernels.metal:839-871) and open-addressing hash
 *     (kernels.metal:786-837) restated as a slot array keyed by key - min;
 *   - star join -> SUM -> top-K: executor/fast_path/mod.rs:58-141 with the
 *     comparator of kernels.metal:1102-1118 extended to a total order
 *     (sum desc, tie columns, parent key asc), computed by full sort.
 * Every function states which closed form or fixture pins it in
 * tests/test_oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/mettle_hip.h"

/* ---------------------------------------------------------------------- */
static int64_t orc_key(const void *col, int32_t type, uint64_t row) {
  switch (type) {
    case MTL_BIGINT:
    case MTL_TIMESTAMP:
    case MTL_DECIMAL:
      return ((const int64_t *)col)[row];
    case MTL_DICT8:
      return (int64_t)((const uint8_t *)col)[row];
    default:
      return (int64_t)((const int32_t *)col)[row];
  }
}

static int cmp3_i64(int64_t a, int64_t b) { return a < b ? -1 : (a > b ? 1 : 0); }

/* DuckDB float order: NaN == NaN, NaN > everything else. */
static int cmp3_f64(double a, double b) {
  int an = isnan(a), bn = isnan(b);
  if (an || bn) return an && bn ? 0 : (an ? 1 : -1);
  return a < b ? -1 : (a > b ? 1 : 0);
}

static int op_holds(int c, int32_t op) {
  switch (op) {
    case MTL_CMP_EQ: return c == 0;
    case MTL_CMP_LT: return c < 0;
    case MTL_CMP_LE: return c <= 0;
    case MTL_CMP_GT: return c > 0;
    default: return c >= 0;
  }
}

/* compare_*_scalar (kernels.metal:70-173) for one row. */
int orc_pred_row(const mtl_pred *p, uint64_t row) {
  int c;
  double s;
  memcpy(&s, &p->scalar, 8);
  switch (p->type) {
    case MTL_FLOAT:
      c = cmp3_f64((double)((const float *)p->column)[row], (double)(float)s);
      break;
    case MTL_DOUBLE:
      c = cmp3_f64(((const double *)p->column)[row], s);
      break;
    case MTL_INTEGER:
    case MTL_DATE:
    case MTL_DICT32:
      /* DuckDB compares an INTEGER column with a BIGINT literal in BIGINT:
       * the literal is never truncated to 32 bits */
      c = cmp3_i64(((const int32_t *)p->column)[row], p->scalar);
      break;
    default:
      c = cmp3_i64(orc_key(p->column, p->type, row), p->scalar);
      break;
  }
  return op_holds(c, p->op);
}

/* Predicate conjunction -> ascending row ids.  Returns the count. */
uint64_t orc_filter(const mtl_pred *preds, uint32_t n_preds, uint64_t n_rows,
                    const uint32_t *in_sel, uint64_t in_n, uint32_t *out_sel) {
  uint64_t n = in_sel ? in_n : n_rows, out = 0;
  for (uint64_t k = 0; k < n; ++k) {
    uint64_t row = in_sel ? in_sel[k] : k;
    int ok = 1;
    for (uint32_t p = 0; p < n_preds && ok; ++p) ok = orc_pred_row(&preds[p], row);
    if (ok) out_sel[out++] = (uint32_t)row;
  }
  return out;
}

/* project_*_scalar op codes (kernels.metal:4-68); exact integers. */
int32_t orc_project(const void *col, int32_t type, const uint32_t *sel, uint64_t n,
                    int32_t op, int64_t scalar, void *out, uint32_t *overflow) {
  *overflow = 0;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t row = sel ? sel[i] : i;
    if (type == MTL_FLOAT) {
      double sd;
      memcpy(&sd, &scalar, 8);
      float x = ((const float *)col)[row], s = (float)sd, r;
      r = op == 0 ? x + s : op == 1 ? x * s : op == 2 ? x - s : s - x;
      ((float *)out)[i] = r;
    } else if (type == MTL_DOUBLE) {
      double x = ((const double *)col)[row], s, r;
      memcpy(&s, &scalar, 8);
      r = op == 0 ? x + s : op == 1 ? x * s : op == 2 ? x - s : s - x;
      ((double *)out)[i] = r;
    } else if (type == MTL_INTEGER || type == MTL_DATE) {
      int64_t x = ((const int32_t *)col)[row], s = (int32_t)scalar, r;
      r = op == 0 ? x + s : op == 1 ? x * s : op == 2 ? x - s : s - x;
      if (r < INT32_MIN || r > INT32_MAX) *overflow = 1;
      ((int32_t *)out)[i] = (int32_t)r;
    } else {
      __int128 x = ((const int64_t *)col)[row], s = scalar, r;
      r = op == 0 ? x + s : op == 1 ? x * s : op == 2 ? x - s : s - x;
      if (r < INT64_MIN || r > INT64_MAX) *overflow = 1;
      ((int64_t *)out)[i] = (int64_t)r;
    }
  }
  return MTL_OK;
}

/* Inner PK/FK equi-join restated as a slot array over the build key span
 * (dense_lookup_build/probe, kernels.metal:839-871).  Emits (probe, build)
 * row pairs in ascending probe order.  Returns the pair count, or -1 if a
 * build key repeats (the host falls back), -2 if the span is too large. */
int64_t orc_join(const void *bkeys, int32_t btype, const uint32_t *bsel, uint64_t bn,
                 const void *pkeys, int32_t ptype, const uint32_t *psel, uint64_t pn,
                 uint32_t *out_probe, uint32_t *out_build) {
  if (bn == 0) return 0;
  int64_t mn = INT64_MAX, mx = INT64_MIN;
  for (uint64_t i = 0; i < bn; ++i) {
    int64_t k = orc_key(bkeys, btype, bsel ? bsel[i] : i);
    if (k < mn) mn = k;
    if (k > mx) mx = k;
  }
  uint64_t span = (uint64_t)mx - (uint64_t)mn + 1;
  if (span > (1ull << 31)) return -2;
  int64_t *slot = (int64_t *)malloc(span * sizeof(int64_t));
  if (!slot) return -2;
  for (uint64_t i = 0; i < span; ++i) slot[i] = -1;
  for (uint64_t i = 0; i < bn; ++i) {
    uint64_t row = bsel ? bsel[i] : i;
    uint64_t s = (uint64_t)orc_key(bkeys, btype, row) - (uint64_t)mn;
    if (slot[s] != -1) {
      free(slot);
      return -1;
    }
    slot[s] = (int64_t)row;
  }
  int64_t out = 0;
  for (uint64_t i = 0; i < pn; ++i) {
    uint64_t row = psel ? psel[i] : i;
    int64_t k = orc_key(pkeys, ptype, row);
    if (k < mn || k > mx) continue;
    int64_t b = slot[(uint64_t)k - (uint64_t)mn];
    if (b < 0) continue;
    out_probe[out] = (uint32_t)row;
    out_build[out] = (uint32_t)b;
    ++out;
  }
  free(slot);
  return out;
}

/* ---------------------------------------------------------------------- */
/* Star join -> SUM(expr) per parent row -> top-K.                         */
/* ---------------------------------------------------------------------- */
static double orc_f64(const void *col, int32_t type, uint64_t row) {
  switch (type) {
    case MTL_FLOAT: return (double)((const float *)col)[row];
    case MTL_DOUBLE: return ((const double *)col)[row];
    default: return (double)orc_key(col, type, row);
  }
}

/* Evaluates the postfix program; int64 mode flags overflow. */
static int64_t orc_expr_i64(const mtl_expr *e, uint64_t row, int *ovf) {
  __int128 st[16];
  int sp = 0;
  for (uint32_t i = 0; i < e->n_ops; ++i) {
    int op = e->op[i];
    if (op == MTL_EXPR_COL) st[sp++] = orc_key(e->col[e->arg[i]], e->col_type[e->arg[i]], row);
    else if (op == MTL_EXPR_CONST) st[sp++] = e->arg[i];
    else {
      __int128 b = st[--sp], a = st[--sp], r;
      r = op == MTL_EXPR_ADD ? a + b : op == MTL_EXPR_SUB ? a - b : a * b;
      if (r < INT64_MIN || r > INT64_MAX) *ovf = 1;
      st[sp++] = (int64_t)r;
    }
  }
  return sp ? (int64_t)st[sp - 1] : 0;
}

static double orc_expr_f64(const mtl_expr *e, uint64_t row) {
  double st[16];
  int sp = 0;
  for (uint32_t i = 0; i < e->n_ops; ++i) {
    int op = e->op[i];
    double a, b, c;
    if (op == MTL_EXPR_COL) st[sp++] = orc_f64(e->col[e->arg[i]], e->col_type[e->arg[i]], row);
    else if (op == MTL_EXPR_CONST) memcpy(&st[sp++], &e->arg[i], 8);
    else {
      b = st[--sp];
      a = st[--sp];
      c = op == MTL_EXPR_ADD ? a + b : op == MTL_EXPR_SUB ? a - b : a * b;
      st[sp++] = c;
    }
  }
  return sp ? st[sp - 1] : 0.0;
}

typedef struct {
  uint64_t slot;
  uint32_t row;
  uint32_t count;
  __int128 isum;
  double fsum;
} orc_group;

static const mtl_star_desc *g_desc;

static int orc_tie_cmp(const void *col, int32_t type, uint32_t a, uint32_t b) {
  if (type == MTL_FLOAT || type == MTL_DOUBLE)
    return cmp3_f64(orc_f64(col, type, a), orc_f64(col, type, b));
  return cmp3_i64(orc_key(col, type, a), orc_key(col, type, b));
}

static int orc_group_cmp(const void *pa, const void *pb) {
  const orc_group *a = (const orc_group *)pa, *b = (const orc_group *)pb;
  int c;
  if (g_desc->agg.is_double) c = -cmp3_f64(a->fsum, b->fsum);
  else c = a->isum > b->isum ? -1 : (a->isum < b->isum ? 1 : 0);
  if (c) return c;
  for (uint32_t t = 0; t < g_desc->n_ties; ++t) {
    c = orc_tie_cmp(g_desc->tie_col[t], g_desc->tie_type[t], a->row, b->row);
    if (g_desc->tie_desc[t]) c = -c;
    if (c) return c;
  }
  return a->slot < b->slot ? -1 : (a->slot > b->slot ? 1 : 0);
}

/* Returns MTL_OK, MTL_UNSUPPORTED (duplicate join key or int64 sum
 * overflow), MTL_ERR_OVERFLOW (expression overflow) or
 * MTL_ERR_INVALID_ARGUMENT (key outside its stated range). */
int32_t orc_star_topk(const mtl_star_desc *d, mtl_star_row *out, uint32_t *out_n) {
  uint8_t *dim_ok = NULL;
  uint64_t dspan = 0, pspan = (uint64_t)d->p_key_max - (uint64_t)d->p_key_min + 1;
  int32_t status = MTL_OK;
  *out_n = 0;
  if (d->d_rows) {
    dspan = (uint64_t)d->d_key_max - (uint64_t)d->d_key_min + 1;
    dim_ok = (uint8_t *)calloc(dspan, 1);
    for (uint64_t i = 0; i < d->d_rows; ++i) {
      int ok = 1;
      for (uint32_t p = 0; p < d->d_n_preds && ok; ++p) ok = orc_pred_row(&d->d_preds[p], i);
      if (!ok) continue;
      int64_t k = orc_key(d->d_key, d->d_key_type, i);
      if (k < d->d_key_min || k > d->d_key_max) { status = MTL_ERR_INVALID_ARGUMENT; goto done; }
      if (dim_ok[k - d->d_key_min]) { status = MTL_UNSUPPORTED; goto done; }
      dim_ok[k - d->d_key_min] = 1;
    }
  }
  int64_t *prow = (int64_t *)malloc(pspan * sizeof(int64_t));
  for (uint64_t i = 0; i < pspan; ++i) prow[i] = -1;
  for (uint64_t i = 0; i < d->p_rows; ++i) {
    int ok = 1;
    for (uint32_t p = 0; p < d->p_n_preds && ok; ++p) ok = orc_pred_row(&d->p_preds[p], i);
    if (!ok) continue;
    if (d->d_rows) {
      int64_t fk = orc_key(d->p_fk, d->p_fk_type, i);
      if (fk < d->d_key_min || fk > d->d_key_max || !dim_ok[fk - d->d_key_min]) continue;
    }
    int64_t k = orc_key(d->p_key, d->p_key_type, i);
    if (k < d->p_key_min || k > d->p_key_max) { status = MTL_ERR_INVALID_ARGUMENT; free(prow); goto done; }
    if (prow[k - d->p_key_min] != -1) { status = MTL_UNSUPPORTED; free(prow); goto done; }
    prow[k - d->p_key_min] = (int64_t)i;
  }
  {
    __int128 *isum = (__int128 *)calloc(pspan, sizeof(__int128));
    double *fsum = (double *)calloc(pspan, sizeof(double));
    double *fcomp = (double *)calloc(pspan, sizeof(double));
    uint32_t *cnt = (uint32_t *)calloc(pspan, sizeof(uint32_t));
    int ovf = 0;
    for (uint64_t i = 0; i < d->f_rows; ++i) {
      int64_t fk = orc_key(d->f_fk, d->f_fk_type, i);
      if (fk < d->p_key_min || fk > d->p_key_max) continue;
      uint64_t s = (uint64_t)fk - (uint64_t)d->p_key_min;
      if (prow[s] < 0) continue;
      int ok = 1;
      for (uint32_t p = 0; p < d->f_n_preds && ok; ++p) ok = orc_pred_row(&d->f_preds[p], i);
      if (!ok) continue;
      if (d->agg.is_double) {
        /* Neumaier compensated sum: the accurate reference for f64 SUM */
        double v = orc_expr_f64(&d->agg, i), t = fsum[s] + v;
        if (fabs(fsum[s]) >= fabs(v)) fcomp[s] += (fsum[s] - t) + v;
        else fcomp[s] += (v - t) + fsum[s];
        fsum[s] = t;
      } else {
        isum[s] += orc_expr_i64(&d->agg, i, &ovf);
      }
      cnt[s]++;
    }
    uint64_t ng = 0;
    for (uint64_t s = 0; s < pspan; ++s) ng += cnt[s] > 0;
    orc_group *g = (orc_group *)malloc((ng ? ng : 1) * sizeof(orc_group));
    uint64_t j = 0;
    for (uint64_t s = 0; s < pspan; ++s) {
      if (!cnt[s]) continue;
      g[j].slot = s;
      g[j].row = (uint32_t)prow[s];
      g[j].count = cnt[s];
      g[j].isum = isum[s];
      g[j].fsum = fsum[s] + fcomp[s];
      if (!d->agg.is_double && (isum[s] < INT64_MIN || isum[s] > INT64_MAX)) status = MTL_UNSUPPORTED;
      ++j;
    }
    if (ovf) status = MTL_ERR_OVERFLOW;
    if (status == MTL_OK) {
      g_desc = d;
      qsort(g, ng, sizeof(orc_group), orc_group_cmp);
      uint32_t k = (uint32_t)(ng < d->k ? ng : d->k);
      for (uint32_t i = 0; i < k; ++i) {
        out[i].p_row = g[i].row;
        out[i].count = g[i].count;
        out[i].sum_i64 = d->agg.is_double ? 0 : (int64_t)g[i].isum;
        out[i].sum_f64 = d->agg.is_double ? g[i].fsum : (double)(int64_t)g[i].isum;
      }
      *out_n = k;
    }
    free(g);
    free(isum);
    free(fsum);
    free(fcomp);
    free(cnt);
  }
  free(prow);
done:
  free(dim_ok);
  return status;
}

/* ---------------------------------------------------------------------- */
/* Grouped aggregate: sort rows by key tuple, aggregate runs.              */
/* Restates grouped_aggregate_exact_gv (aggregate_ops.rs:101-457) and the  */
/* CPU HashMap path (aggregate.rs:128-216) with DuckDB result types:       */
/* COUNT -> BIGINT, SUM(int) -> HUGEINT, SUM/AVG(float) -> DOUBLE          */
/* (compensated sum), AVG(int) -> hugeint / count in long double.          */
/* ---------------------------------------------------------------------- */
static const mtl_group_desc *g_gdesc;

static int orc_row_key_cmp(const void *pa, const void *pb) {
  uint32_t a = *(const uint32_t *)pa, b = *(const uint32_t *)pb;
  for (uint32_t i = 0; i < g_gdesc->n_keys; ++i) {
    int c = cmp3_i64(orc_key(g_gdesc->key_col[i], g_gdesc->key_type[i], a),
                     orc_key(g_gdesc->key_col[i], g_gdesc->key_type[i], b));
    if (c) return c;
  }
  return a < b ? -1 : (a > b ? 1 : 0);
}

/* Returns the number of groups written (ascending by key tuple), or -1 if
 * more than `cap` groups exist. */
int64_t orc_group_agg(const mtl_group_desc *d, int64_t *keys_out, void *const *agg_out,
                      uint64_t cap) {
  uint64_t n = d->sel ? d->sel_capacity : d->n_rows;
  uint32_t *rows = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
  for (uint64_t i = 0; i < n; ++i) rows[i] = d->sel ? d->sel[i] : (uint32_t)i;
  g_gdesc = d;
  qsort(rows, n, sizeof(uint32_t), orc_row_key_cmp);
  int64_t ng = 0;
  uint64_t i = 0;
  while (i < n) {
    uint64_t j = i;
    while (j < n) {
      int same = 1;
      for (uint32_t k = 0; k < d->n_keys && same; ++k)
        same = orc_key(d->key_col[k], d->key_type[k], rows[j]) ==
               orc_key(d->key_col[k], d->key_type[k], rows[i]);
      if (!same) break;
      ++j;
    }
    if ((uint64_t)ng >= cap) {
      free(rows);
      return -1;
    }
    for (uint32_t k = 0; k < d->n_keys; ++k)
      keys_out[ng * d->n_keys + k] = orc_key(d->key_col[k], d->key_type[k], rows[i]);
    int64_t cnt = (int64_t)(j - i);
    for (uint32_t a = 0; a < d->n_aggs; ++a) {
      const mtl_agg_spec *sp = &d->aggs[a];
      int isf = sp->type == MTL_FLOAT || sp->type == MTL_DOUBLE;
      if (sp->func == MTL_AGG_COUNT) {
        ((int64_t *)agg_out[a])[ng] = cnt;
      } else if (sp->func == MTL_AGG_SUM || sp->func == MTL_AGG_AVG) {
        if (isf) {
          double s = 0, c = 0;
          for (uint64_t r = i; r < j; ++r) {
            double v = orc_f64(sp->values, sp->type, rows[r]), t = s + v;
            if (fabs(s) >= fabs(v)) c += (s - t) + v;
            else c += (v - t) + s;
            s = t;
          }
          s += c;
          ((double *)agg_out[a])[ng] = sp->func == MTL_AGG_SUM ? s : s / (double)cnt;
        } else {
          __int128 s = 0;
          for (uint64_t r = i; r < j; ++r) s += orc_key(sp->values, sp->type, rows[r]);
          if (sp->func == MTL_AGG_SUM) {
            ((uint64_t *)agg_out[a])[2 * ng] = (uint64_t)s;
            ((int64_t *)agg_out[a])[2 * ng + 1] = (int64_t)(s >> 64);
          } else {
            /* avg.cpp GetAverageDivident: count * 10^scale for DECIMAL */
            long double div = (long double)cnt;
            if (sp->type == MTL_DECIMAL) {
              double p10 = 1.0;
              for (int32_t q = 0; q < sp->scale; ++q) p10 *= 10.0;
              div *= (long double)p10;
            }
            ((double *)agg_out[a])[ng] = (double)((long double)s / div);
          }
        }
      } else {
        int mx = sp->func == MTL_AGG_MAX;
        if (isf) {
          double best = orc_f64(sp->values, sp->type, rows[i]);
          for (uint64_t r = i + 1; r < j; ++r) {
            double v = orc_f64(sp->values, sp->type, rows[r]);
            int c = cmp3_f64(v, best);
            if (mx ? c > 0 : c < 0) best = v;
          }
          ((double *)agg_out[a])[ng] = best;
        } else {
          int64_t best = orc_key(sp->values, sp->type, rows[i]);
          for (uint64_t r = i + 1; r < j; ++r) {
            int64_t v = orc_key(sp->values, sp->type, rows[r]);
            if (mx ? v > best : v < best) best = v;
          }
          ((int64_t *)agg_out[a])[ng] = best;
        }
      }
    }
    ++ng;
    i = j;
  }
  free(rows);
  return ng;
}

/* Days since 1970-01-01 for a civil date (proleptic Gregorian), the closed
 * form of parse_date_to_days_since_epoch (src/plan/parser.rs:426-449). */
int32_t orc_date_days(int32_t y, int32_t m, int32_t d) {
  int32_t yy = y - (m <= 2);
  int32_t era = (yy >= 0 ? yy : yy - 399) / 400;
  int32_t yoe = yy - era * 400;
  int32_t mp = m + (m > 2 ? -3 : 9);
  int32_t doy = (153 * mp + 2) / 5 + d - 1;
  int32_t doe = yoe * 365 + yoe / 4 - yoe / 100 + doy;
  return era * 146097 + doe - 719468;
}
