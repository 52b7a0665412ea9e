// SQL lowering: the admitted subset of SELECT is parsed, bound against a
// catalog of device-resident tables, and lowered onto the operators of this
// library; anything outside the subset returns MTL_UNSUPPORTED so the host
// runs DuckDB's CPU path.
//
// Reference: try_build_gpu_plan (parser.rs:20-120) admits one SELECT with
// comma or INNER JOIN sources, AND-ed `col op literal` / BETWEEN filters,
// equi-join conditions, GROUP BY columns, SUM/COUNT/MIN/MAX/AVG, ORDER BY
// and LIMIT (parse_where_clause :261-341, extract_literal :364-400, date
// literals :403-431); native_gpu_shape / try_execute (executor/mod.rs:44-138)
// route the plan; the Q3 fast path (executor/fast_path/mod.rs:58-141)
// handles the star join.  The three shapes lowered here:
//   scan  : one table, filters, column / `col op literal` projections,
//           ORDER BY one column, LIMIT       -> mtl_filter, mtl_sort,
//                                               mtl_gather, mtl_project
//   group : one table, GROUP BY integer columns, aggregates over columns,
//           optional ORDER BY one aggregate + LIMIT -> mtl_group_agg, topk
//   star  : fact -> parent (-> dimension) equi-joins on unique keys, GROUP
//           BY parent columns including its key, SUM(expression over fact
//           columns), ORDER BY the sum then parent columns, LIMIT
//                                            -> mtl_star_topk
// Unlike the reference, literals keep the column's exact type (DATE days,
// TIMESTAMP microseconds, DECIMAL scaled integers, dictionary codes), and
// results carry DuckDB's types (exact integer sums, no float narrowing).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "filter.hpp"
#include "join.hpp"
#include "mtl_internal.hpp"
#include "star.hpp"

// ---------------------------------------------------------------------------
// Results (host-side materialised columns)
// ---------------------------------------------------------------------------
struct mtl_result {
  struct Column {
    std::string name;
    int32_t type;
    int32_t scale = 0;  // DECIMAL scale
    std::vector<uint8_t> data;
    std::vector<std::string> dict;  // dictionary columns: the value of each code
    std::vector<uint8_t> valid;     // one byte per row, 1 = valid; empty: no NULLs
  };
  std::vector<Column> cols;
  uint64_t rows = 0;
  std::string shape;
  uint64_t scratch = 0;  // planning only: device bytes the run may take
};

namespace mtl {
uint64_t group_scratch_bytes(const mtl_group_desc& d);  // group.hip
bool group_is_dense(const mtl_group_desc& d);           // group.hip
uint64_t query_scratch(const mtl_query& q);             // serve.hip
namespace sql {

// ---------------------------------------------------------------------------
// Lexer
// ---------------------------------------------------------------------------
struct Tok {
  enum Kind { ID, NUM, STR, OP, END } k;
  std::string s;
};

static std::string lower(std::string s) {
  for (auto& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

static bool lex(const std::string& in, std::vector<Tok>& out, std::string& err) {
  size_t i = 0;
  while (i < in.size()) {
    const char c = in[i];
    if (std::isspace((unsigned char)c)) {
      ++i;
    } else if (std::isalpha((unsigned char)c) || c == '_') {
      size_t j = i;
      while (j < in.size() && (std::isalnum((unsigned char)in[j]) || in[j] == '_')) ++j;
      out.push_back({Tok::ID, lower(in.substr(i, j - i))});
      i = j;
    } else if (c == '"') {
      size_t j = in.find('"', i + 1);
      if (j == std::string::npos) return err = "unterminated identifier", false;
      out.push_back({Tok::ID, lower(in.substr(i + 1, j - i - 1))});
      i = j + 1;
    } else if (std::isdigit((unsigned char)c) ||
               (c == '.' && i + 1 < in.size() && std::isdigit((unsigned char)in[i + 1]))) {
      size_t j = i;
      while (j < in.size() && (std::isdigit((unsigned char)in[j]) || in[j] == '.')) ++j;
      if (j < in.size() && (in[j] == 'e' || in[j] == 'E')) {
        ++j;
        if (j < in.size() && (in[j] == '+' || in[j] == '-')) ++j;
        while (j < in.size() && std::isdigit((unsigned char)in[j])) ++j;
      }
      out.push_back({Tok::NUM, in.substr(i, j - i)});
      i = j;
    } else if (c == '\'') {
      std::string s;
      size_t j = i + 1;
      while (true) {
        if (j >= in.size()) return err = "unterminated string", false;
        if (in[j] == '\'') {
          if (j + 1 < in.size() && in[j + 1] == '\'') {
            s += '\'';
            j += 2;
            continue;
          }
          break;
        }
        s += in[j++];
      }
      out.push_back({Tok::STR, s});
      i = j + 1;
    } else {
      static const char* two[] = {"<=", ">=", "<>", "!="};
      bool done = false;
      for (const char* t : two)
        if (in.compare(i, 2, t) == 0) {
          out.push_back({Tok::OP, t});
          i += 2;
          done = true;
          break;
        }
      if (done) continue;
      if (std::strchr("=<>+-*/(),.;", c)) {
        out.push_back({Tok::OP, std::string(1, c)});
        ++i;
      } else {
        return err = std::string("unexpected character '") + c + "'", false;
      }
    }
  }
  out.push_back({Tok::END, ""});
  return true;
}

// ---------------------------------------------------------------------------
// AST
// ---------------------------------------------------------------------------
struct Lit {
  enum Kind { INT, FLT, STR, DATE, TS } k = INT;
  int64_t i = 0;
  double f = 0;
  std::string s;
  std::string text;  // as written (numbers), for printing
};

struct ColRef {
  std::string rel;  // table name or alias ("" = unqualified)
  std::string name;
};

struct Expr;
using ExprP = std::shared_ptr<Expr>;
struct Expr {
  enum Kind { COL, LIT, ADD, SUB, MUL, AGG, STAR, CALL } k;
  ColRef col;
  Lit lit;
  std::string fn;  // AGG: sum/count/min/max/avg; CALL: hook name
  ExprP a, b;
};

struct Pred {
  ColRef col;
  int op;  // MTL_CMP_*
  Lit lit;
  bool or_next = false;  // OR-ed with the next predicate (IN lists, OR runs)
  int nulls = 0;         // 1: col IS NULL, 2: col IS NOT NULL (op and lit unused)
};

struct JoinEq {
  ColRef l, r;
};

struct Item {
  ExprP e;
  std::string alias;
};

struct OrderTerm {
  ExprP e;
  bool desc = false;
};

struct Query {
  std::vector<Item> items;
  std::vector<std::pair<std::string, std::string>> from;  // (table, alias)
  std::vector<Pred> preds;
  std::vector<JoinEq> joins;
  std::vector<ColRef> group;
  std::vector<OrderTerm> order;
  int64_t limit = -1;
};

// ---------------------------------------------------------------------------
// Parser (recursive descent over the admitted subset)
// ---------------------------------------------------------------------------
struct Parser {
  const std::vector<Tok>& t;
  size_t p = 0;
  std::string err;
  explicit Parser(const std::vector<Tok>& tt) : t(tt) {}

  const Tok& cur() const { return t[p]; }
  bool is_kw(const char* k) const { return cur().k == Tok::ID && cur().s == k; }
  bool is_op(const char* o) const { return cur().k == Tok::OP && cur().s == o; }
  bool accept_kw(const char* k) {
    if (!is_kw(k)) return false;
    ++p;
    return true;
  }
  bool accept_op(const char* o) {
    if (!is_op(o)) return false;
    ++p;
    return true;
  }
  bool fail(const std::string& m) {
    if (err.empty()) err = m;
    return false;
  }
  bool expect_kw(const char* k) { return accept_kw(k) || fail(std::string("expected ") + k); }
  bool expect_op(const char* o) { return accept_op(o) || fail(std::string("expected '") + o + "'"); }

  static bool reserved(const std::string& s) {
    static const char* kw[] = {"select", "from", "where", "group", "by", "order", "limit",
                               "and", "or", "not", "as", "join", "inner", "on", "asc", "desc",
                               "between", "having", "distinct", "union", "left", "right",
                               "full", "outer", "cross", "with", "offset", "in", "is", "null",
                               nullptr};
    for (int i = 0; kw[i]; ++i)
      if (s == kw[i]) return true;
    return false;
  }

  bool literal(Lit& l) {
    bool neg = false;
    if (is_op("-")) {
      ++p;
      neg = true;
    }
    if (cur().k == Tok::NUM) {
      const std::string s = cur().s;
      ++p;
      l.text = (neg ? "-" : "") + s;
      if (s.find_first_of(".eE") == std::string::npos) {
        errno = 0;
        char* end = nullptr;
        unsigned long long u = std::strtoull(s.c_str(), &end, 10);
        if (errno || *end) return fail("integer literal out of range");
        if (!neg && u > (unsigned long long)INT64_MAX) return fail("integer literal out of range");
        if (neg && u > (unsigned long long)INT64_MAX + 1ull)
          return fail("integer literal out of range");
        l.k = Lit::INT;
        l.i = neg ? (int64_t)(0ull - u) : (int64_t)u;
      } else {
        l.k = Lit::FLT;
        l.f = std::strtod(s.c_str(), nullptr);
        if (neg) l.f = -l.f;
      }
      return true;
    }
    if (neg) return fail("expected a number after '-'");
    if (cur().k == Tok::STR) {
      l.k = Lit::STR;
      l.s = cur().s;
      ++p;
      return true;
    }
    if ((is_kw("date") || is_kw("timestamp")) && t[p + 1].k == Tok::STR) {
      const bool ts = is_kw("timestamp");
      p += 1;
      l.k = ts ? Lit::TS : Lit::DATE;
      l.s = cur().s;
      ++p;
      return true;
    }
    return fail("expected a literal");
  }

  bool colref(ColRef& c) {
    if (cur().k != Tok::ID || reserved(cur().s)) return fail("expected a column");
    std::string a = cur().s;
    ++p;
    if (accept_op(".")) {
      if (cur().k != Tok::ID) return fail("expected a column after '.'");
      c.rel = a;
      c.name = cur().s;
      ++p;
    } else {
      c.name = a;
    }
    return true;
  }

  bool is_literal_start() const {
    if (cur().k == Tok::NUM || cur().k == Tok::STR) return true;
    if (is_op("-") && t[p + 1].k == Tok::NUM) return true;
    if ((is_kw("date") || is_kw("timestamp")) && t[p + 1].k == Tok::STR) return true;
    return false;
  }

  bool factor(ExprP& e) {
    e = std::make_shared<Expr>();
    if (accept_op("(")) {
      if (!expr(e)) return false;
      return expect_op(")");
    }
    if (is_literal_start()) {
      e->k = Expr::LIT;
      return literal(e->lit);
    }
    if (cur().k == Tok::ID && t[p + 1].k == Tok::OP && t[p + 1].s == "(") {
      const std::string fn = cur().s;
      p += 2;
      if (fn == "sum" || fn == "count" || fn == "min" || fn == "max" || fn == "avg") {
        e->k = Expr::AGG;
        e->fn = fn;
        if (fn == "count" && accept_op("*")) {
          e->a = nullptr;
        } else {
          if (accept_kw("distinct")) return fail("DISTINCT aggregates are not supported");
          if (!expr(e->a)) return false;
        }
        return expect_op(")");
      }
      if (fn.rfind("gpu_", 0) == 0) {  // the scalar hooks
        e->k = Expr::CALL;
        e->fn = fn;
        if (!expr(e->a) || !expect_op(",") || !expr(e->b)) return false;
        return expect_op(")");
      }
      return fail("function " + fn + " is not supported");
    }
    e->k = Expr::COL;
    return colref(e->col);
  }

  bool term(ExprP& e) {
    if (!factor(e)) return false;
    while (is_op("*") || is_op("/")) {
      if (is_op("/")) return fail("division is not supported");
      ++p;
      auto n = std::make_shared<Expr>();
      n->k = Expr::MUL;
      n->a = e;
      if (!factor(n->b)) return false;
      e = n;
    }
    return true;
  }

  bool expr(ExprP& e) {
    if (!term(e)) return false;
    while (is_op("+") || is_op("-")) {
      auto n = std::make_shared<Expr>();
      n->k = is_op("+") ? Expr::ADD : Expr::SUB;
      ++p;
      n->a = e;
      if (!term(n->b)) return false;
      e = n;
    }
    return true;
  }

  static int cmp_op(const std::string& s) {
    if (s == "=") return MTL_CMP_EQ;
    if (s == "<") return MTL_CMP_LT;
    if (s == "<=") return MTL_CMP_LE;
    if (s == ">") return MTL_CMP_GT;
    if (s == ">=") return MTL_CMP_GE;
    return -1;
  }
  static int flip(int op) {
    switch (op) {
      case MTL_CMP_LT: return MTL_CMP_GT;
      case MTL_CMP_LE: return MTL_CMP_GE;
      case MTL_CMP_GT: return MTL_CMP_LT;
      case MTL_CMP_GE: return MTL_CMP_LE;
      default: return op;
    }
  }

  // One conjunct of WHERE: col op lit, lit op col, col = col, col BETWEEN,
  // hook call, or a parenthesised conjunction.
  bool condition(Query& q) {
    if (is_op("(")) {
      size_t save = p;
      ++p;
      if (conjunction(q) && accept_op(")")) return true;
      p = save;
      err.clear();
      return fail("unsupported parenthesised condition");
    }
    if (cur().k == Tok::ID && cur().s.rfind("gpu_", 0) == 0 && t[p + 1].s == "(") {
      // gpu_<op>_scalar(col, literal) = the predicate itself
      const std::string fn = cur().s;
      p += 2;
      Pred pr;
      static const std::pair<const char*, int> hooks[] = {
          {"gpu_eq_scalar", MTL_CMP_EQ}, {"gpu_lt_scalar", MTL_CMP_LT},
          {"gpu_le_scalar", MTL_CMP_LE}, {"gpu_gt_scalar", MTL_CMP_GT},
          {"gpu_ge_scalar", MTL_CMP_GE}};
      pr.op = -1;
      for (auto& h : hooks)
        if (fn == h.first) pr.op = h.second;
      if (pr.op < 0) return fail(fn + " is not a predicate hook");
      if (!colref(pr.col) || !expect_op(",") || !literal(pr.lit) || !expect_op(")"))
        return false;
      q.preds.push_back(pr);
      return true;
    }
    if (is_literal_start()) {
      Pred pr;
      if (!literal(pr.lit)) return false;
      if (cur().k != Tok::OP || cmp_op(cur().s) < 0) return fail("expected a comparison");
      pr.op = flip(cmp_op(cur().s));
      ++p;
      if (!colref(pr.col)) return false;
      q.preds.push_back(pr);
      return true;
    }
    ColRef c;
    if (!colref(c)) return false;
    if (accept_kw("is")) {
      Pred pr;
      pr.col = c;
      pr.op = MTL_CMP_EQ;
      pr.nulls = accept_kw("not") ? 2 : 1;
      if (!expect_kw("null")) return false;
      q.preds.push_back(pr);
      return true;
    }
    if (is_kw("not")) return fail("NOT is not supported");
    if (accept_kw("in")) {
      // col IN (v1, ..., vk): k equalities OR-ed together
      if (!expect_op("(")) return false;
      do {
        Pred pr;
        pr.col = c;
        pr.op = MTL_CMP_EQ;
        if (!literal(pr.lit)) return false;
        pr.or_next = true;
        q.preds.push_back(pr);
      } while (accept_op(","));
      if (!expect_op(")")) return false;
      q.preds.back().or_next = false;
      return true;
    }
    if (accept_kw("between")) {
      Pred lo, hi;
      lo.col = hi.col = c;
      lo.op = MTL_CMP_GE;
      hi.op = MTL_CMP_LE;
      if (!literal(lo.lit) || !expect_kw("and") || !literal(hi.lit)) return false;
      q.preds.push_back(lo);
      q.preds.push_back(hi);
      return true;
    }
    if (cur().k != Tok::OP || cmp_op(cur().s) < 0) return fail("expected a comparison");
    const int op = cmp_op(cur().s);
    ++p;
    if (is_literal_start()) {
      Pred pr;
      pr.col = c;
      pr.op = op;
      if (!literal(pr.lit)) return false;
      q.preds.push_back(pr);
      return true;
    }
    ColRef d;
    if (!colref(d)) return false;
    if (op != MTL_CMP_EQ) return fail("only equality between columns (joins) is supported");
    q.joins.push_back({c, d});
    return true;
  }

  // The predicates a condition added form one disjunction: a single
  // predicate, an IN list or a parenthesised OR of such (every predicate but
  // the last OR-ed with the next).
  static bool one_disjunction(const Query& q, size_t first) {
    if (q.preds.size() <= first) return false;
    for (size_t i = first; i + 1 < q.preds.size(); ++i)
      if (!q.preds[i].or_next) return false;
    return true;
  }

  // A term of the WHERE conjunction: a condition, or conditions joined by
  // OR when each is a single comparison or an IN list (OR over a conjunction
  // or over a join condition is not admitted).
  bool term(Query& q, bool& used_or) {
    used_or = false;
    const size_t first = q.preds.size(), joins = q.joins.size();
    if (!condition(q)) return false;
    if (!is_kw("or")) return true;
    if (q.joins.size() != joins || !one_disjunction(q, first))
      return fail("OR over a conjunction or a join condition");
    used_or = true;
    while (accept_kw("or")) {
      q.preds.back().or_next = true;
      const size_t at = q.preds.size();
      if (!condition(q)) return false;
      if (q.joins.size() != joins || !one_disjunction(q, at))
        return fail("OR over a conjunction or a join condition");
    }
    return true;
  }

  // AND binds tighter than OR: an OR outside parentheses is admitted only
  // when it is the whole conjunction (`a OR b`), not beside an AND.
  bool conjunction(Query& q) {
    bool used_or = false, any_or = false;
    size_t terms = 1;
    if (!term(q, used_or)) return false;
    any_or |= used_or;
    while (accept_kw("and")) {
      if (!term(q, used_or)) return false;
      any_or |= used_or;
      ++terms;
    }
    if (is_kw("or") || (any_or && terms > 1))
      return fail("OR beside AND needs parentheses");
    return true;
  }

  bool table_ref(Query& q) {
    if (cur().k != Tok::ID || reserved(cur().s)) return fail("expected a table");
    std::string name = cur().s, alias;
    ++p;
    if (accept_op(".")) {  // schema.table: keep the table
      if (cur().k != Tok::ID) return fail("expected a table");
      name = cur().s;
      ++p;
    }
    if (accept_kw("as")) {
      if (cur().k != Tok::ID) return fail("expected an alias");
      alias = cur().s;
      ++p;
    } else if (cur().k == Tok::ID && !reserved(cur().s)) {
      alias = cur().s;
      ++p;
    }
    q.from.push_back({name, alias});
    return true;
  }

  bool query(Query& q) {
    if (!expect_kw("select")) return false;
    if (is_kw("distinct")) return fail("DISTINCT is not supported");
    do {
      Item it;
      if (accept_op("*")) {
        it.e = std::make_shared<Expr>();
        it.e->k = Expr::STAR;
      } else if (!expr(it.e)) {
        return false;
      }
      if (accept_kw("as")) {
        if (cur().k != Tok::ID) return fail("expected an alias");
        it.alias = cur().s;
        ++p;
      } else if (cur().k == Tok::ID && !reserved(cur().s)) {
        it.alias = cur().s;
        ++p;
      }
      q.items.push_back(it);
    } while (accept_op(","));
    if (!expect_kw("from")) return false;
    if (!table_ref(q)) return false;
    while (true) {
      if (accept_op(",")) {
        if (!table_ref(q)) return false;
      } else if (is_kw("join") || is_kw("inner")) {
        accept_kw("inner");
        if (!expect_kw("join") || !table_ref(q) || !expect_kw("on")) return false;
        ColRef a, b;
        if (!colref(a) || !expect_op("=") || !colref(b)) return false;
        q.joins.push_back({a, b});
        while (accept_kw("and"))
          if (!condition(q)) return false;
      } else {
        break;
      }
    }
    if (accept_kw("where") && !conjunction(q)) return false;
    if (accept_kw("group")) {
      if (!expect_kw("by")) return false;
      do {
        ColRef c;
        if (!colref(c)) return false;
        q.group.push_back(c);
      } while (accept_op(","));
    }
    if (is_kw("having")) return fail("HAVING is not supported");
    if (accept_kw("order")) {
      if (!expect_kw("by")) return false;
      do {
        OrderTerm o;
        if (!expr(o.e)) return false;
        if (accept_kw("desc")) o.desc = true;
        else accept_kw("asc");
        if (is_kw("nulls")) return fail("NULLS FIRST/LAST is not supported");
        q.order.push_back(o);
      } while (accept_op(","));
    }
    if (accept_kw("limit")) {
      if (cur().k != Tok::NUM) return fail("expected a LIMIT count");
      q.limit = std::strtoll(cur().s.c_str(), nullptr, 10);
      ++p;
    }
    if (is_kw("offset")) return fail("OFFSET is not supported");
    accept_op(";");
    if (cur().k != Tok::END) return fail("unexpected '" + cur().s + "'");
    return true;
  }
};

// Days since 1970-01-01 of a proper calendar date (proleptic Gregorian),
// the civil-from-days inverse the planner used (parser.rs:406-431).
static bool parse_date(const std::string& s, int64_t& days) {
  int y, m, d;
  char tail;
  if (std::sscanf(s.c_str(), "%d-%d-%d%c", &y, &m, &d, &tail) != 3) return false;
  static const int mdays[] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
  const bool leap = (y % 4 == 0 && y % 100 != 0) || y % 400 == 0;
  if (m < 1 || m > 12 || d < 1 || d > mdays[m - 1] + (m == 2 && leap)) return false;
  const int64_t yy = y - (m <= 2);
  const int64_t era = (yy >= 0 ? yy : yy - 399) / 400;
  const int64_t yoe = yy - era * 400;
  const int64_t mp = m + (m > 2 ? -3 : 9);
  const int64_t doy = (153 * mp + 2) / 5 + d - 1;
  const int64_t doe = yoe * 365 + yoe / 4 - yoe / 100 + doy;
  days = era * 146097 + doe - 719468;
  return true;
}

static bool parse_timestamp(const std::string& s, int64_t& us) {
  int64_t days;
  const std::string date = s.substr(0, s.find_first_of(" T"));
  if (!parse_date(date, days)) return false;
  int hh = 0, mm = 0;
  double ss = 0;
  if (date.size() < s.size()) {
    if (std::sscanf(s.c_str() + date.size() + 1, "%d:%d:%lf", &hh, &mm, &ss) < 2) return false;
    if (hh < 0 || hh > 23 || mm < 0 || mm > 59 || ss < 0 || ss >= 60) return false;
  }
  us = ((days * 24 + hh) * 60 + mm) * 60 * 1000000ll + (int64_t)std::llround(ss * 1e6);
  return true;
}

// ---------------------------------------------------------------------------
// Catalog of device-resident tables
// ---------------------------------------------------------------------------
struct CatCol {
  std::string name;
  int32_t type = 0;
  int32_t scale = 0;  // DECIMAL
  const void* dev = nullptr;
  int64_t min = 0, max = 0;
  bool unique = false;
  std::vector<std::string> dict;  // dictionary columns: code -> value
  // NULLs: one byte per row, 1 = valid (null: no NULLs).  The value bytes of
  // a NULL row are zero, and min/max bound the valid values.
  const uint8_t* valid = nullptr;
};
struct CatTable {
  std::string name;
  uint64_t rows = 0;
  std::vector<CatCol> cols;
  const CatCol* find(const std::string& n) const {
    for (auto& c : cols)
      if (c.name == n) return &c;
    return nullptr;
  }
};

static std::mutex g_cat_mu;
static std::map<std::string, CatTable>& catalog() {
  static std::map<std::string, CatTable> c;
  return c;
}

// ---------------------------------------------------------------------------
// Binding
// ---------------------------------------------------------------------------
struct Bound {
  std::vector<const CatTable*> tabs;
  std::vector<std::string> names;  // alias or table name, per source
  std::string err;

  int source_of(const std::string& rel) const {
    for (size_t i = 0; i < tabs.size(); ++i)
      if (names[i] == rel || tabs[i]->name == rel) return (int)i;
    return -1;
  }
  // resolve a column to (source, column); unqualified must be unambiguous
  bool resolve(const ColRef& c, int& src, const CatCol*& col) {
    src = -1;
    col = nullptr;
    if (!c.rel.empty()) {
      src = source_of(c.rel);
      if (src < 0) return err = "unknown table " + c.rel, false;
      col = tabs[src]->find(c.name);
      if (!col) return err = "unknown column " + c.rel + "." + c.name, false;
      return true;
    }
    for (size_t i = 0; i < tabs.size(); ++i)
      if (auto* cc = tabs[i]->find(c.name)) {
        if (col) return err = "ambiguous column " + c.name, false;
        src = (int)i;
        col = cc;
      }
    if (!col) return err = "unknown column " + c.name, false;
    return true;
  }
};

static bool pow10_scale(int64_t v, int sc, int64_t& out) {
  out = v;
  for (int i = 0; i < sc; ++i)
    if (__builtin_mul_overflow(out, (int64_t)10, &out)) return false;
  return true;
}

// Literal -> the int64 bits mtl_pred expects for the column.  Integer
// literals compare exactly (DuckDB widens, never truncates).
// The literal of `col op lit` in the column's representation.  A string
// compared with a dictionary column is its code; with a sorted dictionary
// (codes order like the strings) <, <=, >, >= become code comparisons and
// op may change (s <= 'x' -> code < first code above 'x').
static bool literal_bits(const Lit& l, const CatCol& c, int& op, int64_t& bits, std::string& err) {
  switch (c.type) {
    case MTL_FLOAT:
    case MTL_DOUBLE: {
      double d;
      if (l.k == Lit::INT) d = (double)l.i;
      else if (l.k == Lit::FLT) d = l.f;
      else return err = "non-numeric literal for " + c.name, false;
      std::memcpy(&bits, &d, 8);
      return true;
    }
    case MTL_DATE:
      if (l.k == Lit::DATE || l.k == Lit::STR) {
        if (!parse_date(l.s, bits)) return err = "invalid date '" + l.s + "'", false;
        return true;
      }
      return err = "DATE column compared with a non-date literal", false;
    case MTL_TIMESTAMP:
      if (l.k == Lit::TS || l.k == Lit::DATE || l.k == Lit::STR) {
        if (!parse_timestamp(l.s, bits)) return err = "invalid timestamp '" + l.s + "'", false;
        return true;
      }
      return err = "TIMESTAMP column compared with a non-timestamp literal", false;
    case MTL_DECIMAL: {
      if (l.k == Lit::INT) {
        if (!pow10_scale(l.i, c.scale, bits)) return err = "decimal literal out of range", false;
        return true;
      }
      if (l.k == Lit::FLT) {  // exact only when the literal has <= scale decimals
        const double scaled = l.f * std::pow(10.0, c.scale);
        if (scaled != std::nearbyint(scaled) || std::fabs(scaled) > 9.0e15)
          return err = "decimal literal needs more scale than the column", false;
        bits = (int64_t)std::nearbyint(scaled);
        return true;
      }
      return err = "non-numeric literal for a DECIMAL column", false;
    }
    case MTL_DICT8:
    case MTL_DICT32:
      if (l.k != Lit::STR) return err = "dictionary column compared with a non-string", false;
      if (op != MTL_CMP_EQ) {
        if (!std::is_sorted(c.dict.begin(), c.dict.end()))
          return err = "only '=' on a dictionary column whose dictionary is not sorted", false;
        const int64_t lb = std::lower_bound(c.dict.begin(), c.dict.end(), l.s) - c.dict.begin();
        const int64_t ub = std::upper_bound(c.dict.begin(), c.dict.end(), l.s) - c.dict.begin();
        switch (op) {
          case MTL_CMP_LT: bits = lb; break;                     // code < lb
          case MTL_CMP_LE: op = MTL_CMP_LT; bits = ub; break;    // code < ub
          case MTL_CMP_GT: op = MTL_CMP_GE; bits = ub; break;    // code >= ub
          default: bits = lb; break;                             // GE: code >= lb
        }
        return true;
      }
      for (size_t i = 0; i < c.dict.size(); ++i)
        if (c.dict[i] == l.s) {
          bits = (int64_t)i;
          return true;
        }
      bits = -1;  // matches no code (literals compare exactly as integers)
      return true;
    default:  // INTEGER, BIGINT
      if (l.k == Lit::INT) {
        bits = l.i;
        return true;
      }
      return err = "integer column compared with a non-integer literal", false;
  }
}

// ---------------------------------------------------------------------------
// Execution helpers
// ---------------------------------------------------------------------------
// The WHERE predicates bound per source: preds[src] with or_next[src] (bit i:
// predicate i is OR-ed with predicate i + 1).  A run of OR-ed predicates must
// stay within one table.
// A predicate no row passes: col < the smallest value of its type.
static mtl_pred never_pred(const CatCol& c) {
  mtl_pred mp{};
  mp.column = c.dev;
  mp.type = c.type;
  mp.op = MTL_CMP_LT;
  if (c.type == MTL_FLOAT || c.type == MTL_DOUBLE) {
    const double d = -HUGE_VAL;
    std::memcpy(&mp.scalar, &d, 8);
  } else if (c.type == MTL_DICT8 || c.type == MTL_DICT32) {
    mp.scalar = 0;
  } else {
    mp.scalar = INT64_MIN;
  }
  return mp;
}

// valid = 1 (or 0): the rows where a column with NULLs is not NULL (is NULL).
static mtl_pred valid_pred(const CatCol& c, int64_t v = 1) {
  mtl_pred mp{};
  mp.column = c.valid;
  mp.type = MTL_DICT8;
  mp.op = MTL_CMP_EQ;
  mp.scalar = v;
  return mp;
}

// WHERE -> per-table predicate lists and OR masks.  NULL semantics: a
// comparison with NULL is not true, so a run of OR-ed comparisons on one
// column with NULLs is AND-ed with that column's valid = 1 (a run over
// several columns, any with NULLs, would need OR of ANDs: refused);
// IS [NOT] NULL tests the valid bytes.
static int32_t bind_preds(const Query& q, Bound& b, std::vector<std::vector<mtl_pred>>& preds,
                          std::vector<uint32_t>& or_next) {
  preds.assign(b.tabs.size(), {});
  or_next.assign(b.tabs.size(), 0u);
  auto push = [&](int src, const mtl_pred& mp, bool ored) -> int32_t {
    if (preds[src].size() >= MTL_MAX_PREDS)
      return set_error(MTL_UNSUPPORTED, "more than %d predicates on a table", MTL_MAX_PREDS);
    if (ored) or_next[src] |= 1u << preds[src].size();
    preds[src].push_back(mp);
    return MTL_OK;
  };
  int run_src = -1;
  const CatCol* run_col = nullptr;
  bool run_mixed = false, run_null = false;
  for (auto& pr : q.preds) {
    int src;
    const CatCol* c;
    if (!b.resolve(pr.col, src, c)) return set_error(MTL_UNSUPPORTED, "%s", b.err.c_str());
    if (run_src >= 0 && src != run_src)
      return set_error(MTL_UNSUPPORTED, "OR across tables");
    if (pr.nulls) {
      if (run_src >= 0 || pr.or_next) return set_error(MTL_UNSUPPORTED, "IS NULL inside an OR");
      if (c->valid) MTL_TRY(push(src, valid_pred(*c, pr.nulls == 2 ? 1 : 0), false));
      else if (pr.nulls == 1) MTL_TRY(push(src, never_pred(*c), false));
      continue;  // IS NOT NULL of a column without NULLs holds for every row
    }
    mtl_pred mp{};
    mp.column = c->dev;
    mp.type = c->type;
    mp.op = pr.op;
    std::string err;
    if (!literal_bits(pr.lit, *c, mp.op, mp.scalar, err))
      return set_error(MTL_UNSUPPORTED, "%s", err.c_str());
    MTL_TRY(push(src, mp, pr.or_next));
    if (run_src < 0) {
      run_col = c;
      run_mixed = false;
      run_null = false;
    } else if (c != run_col) {
      run_mixed = true;
    }
    run_null |= c->valid != nullptr;
    if (pr.or_next) {
      run_src = src;
      continue;
    }
    if (run_null) {
      if (run_mixed) return set_error(MTL_UNSUPPORTED, "OR over several columns with NULLs");
      MTL_TRY(push(src, valid_pred(*run_col), false));
    }
    run_src = -1;
  }
  return MTL_OK;
}

// Key columns with NULLs: NULL keys match nothing (joins).
static int32_t add_not_null(std::vector<std::vector<mtl_pred>>& preds, int src, const CatCol& c) {
  if (!c.valid) return MTL_OK;
  for (auto& p : preds[src])
    if (p.column == c.valid && p.type == MTL_DICT8 && p.op == MTL_CMP_EQ && p.scalar == 1)
      return MTL_OK;
  if (preds[src].size() >= MTL_MAX_PREDS)
    return set_error(MTL_UNSUPPORTED, "more than %d predicates on a table", MTL_MAX_PREDS);
  preds[src].push_back(valid_pred(c));
  return MTL_OK;
}

static bool any_or(const Query& q) {
  for (auto& pr : q.preds)
    if (pr.or_next) return true;
  return false;
}


// Device buffer charged to the processing budget (MTL_ERR_BUDGET when it
// would exceed it).
struct DevBuf {
  void* p = nullptr;
  uint64_t n = 0;
  ~DevBuf() { dev_free(p, n); }
  int32_t alloc(uint64_t bytes) {
    if (p) dev_free(p, n);  // a buffer reused for a new result
    p = nullptr;
    n = 0;
    MTL_TRY(dev_alloc(bytes, &p, "SQL buffer"));
    n = bytes;
    return MTL_OK;
  }
};

static int32_t download(std::vector<uint8_t>& dst, const void* dev, uint64_t bytes,
                        mtl_stream* s) {
  dst.resize(bytes);
  if (!bytes) return MTL_OK;
  return mtl_memcpy_d2h(dst.data(), dev, bytes, s);
}

static bool is_int_type(int32_t t) {
  return t == MTL_INTEGER || t == MTL_BIGINT || t == MTL_DATE || t == MTL_TIMESTAMP ||
         t == MTL_DECIMAL || t == MTL_DICT8 || t == MTL_DICT32;
}

// gpu_add_scalar / gpu_mul_scalar calls and col op literal -> projection
static bool as_projection(const ExprP& e, ColRef& col, int& op, Lit& lit) {
  if (e->k == Expr::CALL && (e->fn == "gpu_add_scalar" || e->fn == "gpu_mul_scalar")) {
    if (e->a->k != Expr::COL || e->b->k != Expr::LIT) return false;
    col = e->a->col;
    op = e->fn == "gpu_add_scalar" ? MTL_PROJ_ADD : MTL_PROJ_MUL;
    lit = e->b->lit;
    return true;
  }
  if (e->k != Expr::ADD && e->k != Expr::SUB && e->k != Expr::MUL) return false;
  if (e->a->k == Expr::COL && e->b->k == Expr::LIT) {
    col = e->a->col;
    lit = e->b->lit;
    op = e->k == Expr::ADD ? MTL_PROJ_ADD : (e->k == Expr::MUL ? MTL_PROJ_MUL : MTL_PROJ_SUB);
    return true;
  }
  if (e->a->k == Expr::LIT && e->b->k == Expr::COL) {
    col = e->b->col;
    lit = e->a->lit;
    op = e->k == Expr::ADD ? MTL_PROJ_ADD : (e->k == Expr::MUL ? MTL_PROJ_MUL : MTL_PROJ_SUBFROM);
    return true;
  }
  return false;
}

static std::string item_name(const Item& it, size_t i) {
  if (!it.alias.empty()) return it.alias;
  if (it.e->k == Expr::COL) return it.e->col.name;
  return "col" + std::to_string(i);
}

// ---------------------------------------------------------------------------
// Shape 1: scan (filter -> ORDER BY -> LIMIT -> projections)
// ---------------------------------------------------------------------------
static int32_t run_scan(const Query& q, Bound& b, mtl_result& res, mtl_stream* s, bool dry) {
  const CatTable& T = *b.tabs[0];
  std::vector<std::vector<mtl_pred>> tp;
  std::vector<uint32_t> tor;
  MTL_TRY(bind_preds(q, b, tp, tor));
  const std::vector<mtl_pred>& preds = tp[0];
  // outputs
  struct Out {
    std::string name;
    const CatCol* c;
    int op = -1;  // projection op, -1 = plain column
    int64_t scalar = 0;
  };
  std::vector<Out> outs;
  for (size_t i = 0; i < q.items.size(); ++i) {
    const Item& it = q.items[i];
    if (it.e->k == Expr::STAR) {
      for (auto& c : T.cols) outs.push_back({c.name, &c});
      continue;
    }
    Out o;
    o.name = item_name(it, i);
    int src;
    if (it.e->k == Expr::COL) {
      if (!b.resolve(it.e->col, src, o.c)) return set_error(MTL_UNSUPPORTED, "%s", b.err.c_str());
    } else {
      ColRef cr;
      Lit lit;
      if (!as_projection(it.e, cr, o.op, lit))
        return set_error(MTL_UNSUPPORTED, "select item %zu is not a column or col op literal", i);
      if (!b.resolve(cr, src, o.c)) return set_error(MTL_UNSUPPORTED, "%s", b.err.c_str());
      if (o.c->type == MTL_DICT8 || o.c->type == MTL_DICT32 || o.c->type == MTL_TIMESTAMP)
        return set_error(MTL_UNSUPPORTED, "arithmetic on column %s", o.c->name.c_str());
      if (o.c->type == MTL_DATE && o.op != MTL_PROJ_ADD && o.op != MTL_PROJ_SUB)
        return set_error(MTL_UNSUPPORTED, "DATE supports only + and - of days");
      if (is_int_type(o.c->type)) {
        if (lit.k != Lit::INT) return set_error(MTL_UNSUPPORTED, "non-integer literal");
        o.scalar = lit.i;
        if (o.c->type == MTL_DECIMAL && (o.op == MTL_PROJ_ADD || o.op == MTL_PROJ_SUB ||
                                         o.op == MTL_PROJ_SUBFROM) &&
            !pow10_scale(lit.i, o.c->scale, o.scalar))
          return set_error(MTL_UNSUPPORTED, "decimal literal out of range");
      } else {
        double d = lit.k == Lit::INT ? (double)lit.i : lit.f;
        if (lit.k != Lit::INT && lit.k != Lit::FLT)
          return set_error(MTL_UNSUPPORTED, "non-numeric literal");
        std::memcpy(&o.scalar, &d, 8);
      }
    }
    outs.push_back(o);
  }
  const CatCol* okey = nullptr;
  bool odesc = false;
  if (q.order.size() > 1) return set_error(MTL_UNSUPPORTED, "ORDER BY more than one column");
  if (q.order.size() == 1) {
    const ExprP& e = q.order[0].e;
    if (e->k != Expr::COL) return set_error(MTL_UNSUPPORTED, "ORDER BY an expression");
    int src;
    // an output alias of a plain column, or a table column
    for (auto& o : outs)
      if (e->col.rel.empty() && o.name == e->col.name && o.op < 0) okey = o.c;
    if (!okey && !b.resolve(e->col, src, okey))
      return set_error(MTL_UNSUPPORTED, "%s", b.err.c_str());
    odesc = q.order[0].desc;
  }
  res.shape = "scan";
  // selection, count, sort permutation and one output column at a time
  res.scratch = T.rows * (4 + 4 + 8 + 8) + (okey && okey->valid ? T.rows * 8 : 0) + (1u << 20);
  if (dry) return MTL_OK;
  MTL_TRY(resolve_stream(s, &s));  // NULL: the library's default stream
  // filter
  DevBuf sel, cnt, perm, out, ovf;
  MTL_TRY(sel.alloc(std::max<uint64_t>(T.rows, 1) * 4));
  MTL_TRY(cnt.alloc(16));
  MTL_TRY(mtl_filter_or(preds.data(), (uint32_t)preds.size(), tor[0], T.rows,
                        (uint32_t*)sel.p, (uint32_t*)cnt.p, s));
  uint32_t n = 0;
  const uint32_t* rows = (const uint32_t*)sel.p;
  if (okey && okey->valid) {
    // NULLs last (DuckDB's default order): the valid rows sorted, then the
    // NULL rows in selection order
    DevBuf vsel, nsel, vcnt;
    MTL_TRY(vsel.alloc(std::max<uint64_t>(T.rows, 1) * 4));
    MTL_TRY(nsel.alloc(std::max<uint64_t>(T.rows, 1) * 4));
    MTL_TRY(vcnt.alloc(16));
    const mtl_pred pv = valid_pred(*okey, 1), pn = valid_pred(*okey, 0);
    MTL_TRY(mtl_filter(&pv, 1, T.rows, (const uint32_t*)sel.p, (const uint32_t*)cnt.p, T.rows,
                       (uint32_t*)vsel.p, (uint32_t*)vcnt.p, s));
    MTL_TRY(mtl_filter(&pn, 1, T.rows, (const uint32_t*)sel.p, (const uint32_t*)cnt.p, T.rows,
                       (uint32_t*)nsel.p, (uint32_t*)vcnt.p + 1, s));
    uint32_t c2[2] = {0, 0};
    MTL_TRY(mtl_memcpy_d2h(c2, vcnt.p, 8, s));
    if (c2[0] > T.rows || c2[1] > T.rows - c2[0])
      return set_error(MTL_ERR_PROCESSING, "device scan failed");
    MTL_TRY(perm.alloc(std::max<uint64_t>(T.rows, 1) * 4));
    MTL_TRY(mtl_sort(okey->dev, okey->type, (const uint32_t*)vsel.p, (const uint32_t*)vcnt.p,
                     T.rows, odesc ? 1 : 0, (uint32_t*)perm.p, &n, s));
    if (c2[1])
      MTL_HIP(hipMemcpyAsync((uint32_t*)perm.p + n, nsel.p, (uint64_t)c2[1] * 4,
                             hipMemcpyDeviceToDevice, s->stream));
    n += c2[1];
    rows = (const uint32_t*)perm.p;
  } else if (okey) {
    MTL_TRY(perm.alloc(std::max<uint64_t>(T.rows, 1) * 4));
    MTL_TRY(mtl_sort(okey->dev, okey->type, (const uint32_t*)sel.p, (const uint32_t*)cnt.p,
                     T.rows, odesc ? 1 : 0, (uint32_t*)perm.p, &n, s));
    rows = (const uint32_t*)perm.p;
  } else {
    MTL_TRY(mtl_memcpy_d2h(&n, cnt.p, 4, s));
    if (n > T.rows) return set_error(MTL_ERR_PROCESSING, "device scan failed");
  }
  if (q.limit >= 0 && (uint64_t)q.limit < n) n = (uint32_t)q.limit;
  res.rows = n;
  MTL_TRY(out.alloc((uint64_t)std::max<uint32_t>(n, 1) * 8));
  MTL_TRY(ovf.alloc(16));
  for (auto& o : outs) {
    mtl_result::Column col;
    col.name = o.name;
    col.type = o.c->type;
    col.scale = o.c->scale;
    col.dict = o.c->dict;
    const uint32_t w = (uint32_t)type_width(o.c->type);
    if (o.op < 0) {
      MTL_TRY(mtl_gather(o.c->dev, w, rows, nullptr, n, out.p, s));
    } else {
      uint32_t zero = 0;
      MTL_TRY(mtl_memcpy_h2d(ovf.p, &zero, 4, s));
      MTL_TRY(mtl_project(o.c->dev, o.c->type, rows, nullptr, n, o.op, o.scalar, out.p,
                          (uint32_t*)ovf.p, s));
      uint32_t flag = 0;
      MTL_TRY(mtl_memcpy_d2h(&flag, ovf.p, 4, s));
      if (flag) return set_error(MTL_ERR_OVERFLOW, "Overflow in %s of %s",
                                 o.op == MTL_PROJ_MUL ? "multiplication" : "addition",
                                 o.c->name.c_str());
    }
    MTL_TRY(download(col.data, out.p, (uint64_t)n * w, s));
    if (o.c->valid) {  // a projection of NULL is NULL
      if (n) MTL_TRY(mtl_gather(o.c->valid, 1, rows, nullptr, n, out.p, s));
      MTL_TRY(download(col.valid, out.p, n, s));
    }
    res.cols.push_back(std::move(col));
  }
  return MTL_OK;
}

// Output column of a grouped query: a group key or an aggregate.
struct GroupOut {
  std::string name;
  int key = -1, agg = -1;
  // NULLs: nn = the hidden aggregate counting the valid values (-1: none);
  // mode 1 = COUNT(col) (the engine sums an int32 0/1 column), 2 = AVG
  // computed on the host from the engine's SUM and that count
  int nn = -1, mode = 0;
};
// How emit_groups reads the group table back: top-K by one aggregate
// (order_agg >= 0, k = LIMIT), or every group in the table's order
// (key_order 1: ascending by the keys, which a dense table gives; -1: that
// order reversed; 0: as fetched), then LIMIT.
struct GroupEmit {
  const std::vector<const CatCol*>& keys;
  const std::vector<int32_t>& agg_func;
  const std::vector<int32_t>& agg_type;
  const std::vector<int32_t>& agg_scale;
  int order_agg;
  bool odesc;
  int key_order;
  // ORDER BY the aggregate, then these group keys (index, descending)
  std::vector<std::pair<int, bool>> then = {};
  // per group key: the value standing for NULL (INT64_MIN: the key has none)
  std::vector<int64_t> key_null = {};
};

// DuckDB's result type of an aggregate: COUNT BIGINT, integer SUM HUGEINT,
// AVG and float aggregates DOUBLE, DECIMAL SUM/MIN/MAX keeping the scale.
static void agg_out_type(int32_t f, int32_t t, int32_t scale, int32_t& type, int32_t& out_scale) {
  if (f == MTL_AGG_COUNT) type = MTL_BIGINT;
  else if (f == MTL_AGG_SUM && is_int_type(t)) type = MTL_HUGEINT;
  else if (f == MTL_AGG_AVG || t == MTL_FLOAT || t == MTL_DOUBLE) type = MTL_DOUBLE;
  else type = MTL_BIGINT;
  out_scale = 0;
  if ((f == MTL_AGG_SUM || f == MTL_AGG_MIN || f == MTL_AGG_MAX) && t == MTL_DECIMAL)
    out_scale = scale;  // AVG is a DOUBLE: the group engine divided by 10^scale
}

// Three-way comparison of aggregate a between groups i and j of host
// buffers in mtl_groups_fetch layout.
static int agg_cmp(const GroupEmit& ge, const std::vector<std::vector<uint8_t>>& abuf, size_t a,
                   uint64_t w, uint64_t i, uint64_t j) {
  const uint8_t* x = abuf[a].data() + i * w;
  const uint8_t* y = abuf[a].data() + j * w;
  const int32_t f = ge.agg_func[a], t = ge.agg_type[a];
  if (w == 16) {  // int128 {uint64 lower; int64 upper}
    int64_t xh, yh;
    uint64_t xl, yl;
    std::memcpy(&xl, x, 8);
    std::memcpy(&xh, x + 8, 8);
    std::memcpy(&yl, y, 8);
    std::memcpy(&yh, y + 8, 8);
    if (xh != yh) return xh < yh ? -1 : 1;
    return xl == yl ? 0 : (xl < yl ? -1 : 1);
  }
  if (f != MTL_AGG_COUNT && (f == MTL_AGG_AVG || t == MTL_FLOAT || t == MTL_DOUBLE)) {
    double u, v;
    std::memcpy(&u, x, 8);
    std::memcpy(&v, y, 8);
    return u == v ? 0 : (u < v ? -1 : 1);
  }
  int64_t u, v;
  std::memcpy(&u, x, 8);
  std::memcpy(&v, y, 8);
  return u == v ? 0 : (u < v ? -1 : 1);
}

// ORDER BY an aggregate, then group keys, LIMIT k: the device top-K by the
// aggregate takes a wider candidate set (every group tied with the k-th
// lies in it once the last candidate's value differs from the k-th; failing
// that, every group is fetched), which is then ordered by the full ORDER BY
// on the host and cut to k.
template <typename W>
static int32_t topk_then_keys(mtl_groups* g, const mtl_group_desc& d, const GroupEmit& ge,
                              uint32_t k, std::vector<int64_t>& kbuf,
                              std::vector<std::vector<uint8_t>>& abuf, std::vector<void*>& aptr,
                              W agg_width, uint32_t& n_out) {
  const uint32_t a0 = (uint32_t)ge.order_agg;
  const uint64_t w0 = agg_width(a0);
  auto resize = [&](uint64_t cap) {
    kbuf.resize(std::max<uint64_t>(cap, 1) * d.n_keys);
    for (uint32_t a = 0; a < d.n_aggs; ++a) {
      abuf[a].resize(std::max<uint64_t>(cap, 1) * agg_width(a));
      aptr[a] = abuf[a].data();
    }
  };
  const uint32_t kc = std::min<uint32_t>(512, std::max<uint32_t>(2 * k, k + 32));
  uint64_t n = 0;
  bool whole = k > 512;
  if (!whole) {
    resize(kc);
    uint32_t got = 0;
    MTL_TRY(mtl_groups_topk(g, a0, ge.odesc ? 1 : 0, kc, kbuf.data(), aptr.data(), &got));
    n = got;
    whole = got == kc && got > k && agg_cmp(ge, abuf, a0, w0, k - 1, got - 1) == 0;
  }
  if (whole) {
    uint64_t ng = 0;
    MTL_TRY(mtl_groups_count(g, &ng));
    resize(ng);
    MTL_TRY(mtl_groups_fetch(g, kbuf.data(), aptr.data(), ng, &n));
  }
  std::vector<uint64_t> perm(n);
  for (uint64_t i = 0; i < n; ++i) perm[i] = i;
  auto less = [&](uint64_t i, uint64_t j) {
    int c = agg_cmp(ge, abuf, a0, w0, i, j);
    if (c) return ge.odesc ? c > 0 : c < 0;
    for (auto& t : ge.then) {
      const int64_t u = kbuf[i * d.n_keys + t.first], v = kbuf[j * d.n_keys + t.first];
      if (u != v) return t.second ? u > v : u < v;
    }
    for (uint32_t q = 0; q < d.n_keys; ++q) {
      const int64_t u = kbuf[i * d.n_keys + q], v = kbuf[j * d.n_keys + q];
      if (u != v) return u < v;
    }
    return false;
  };
  const uint64_t keep = std::min<uint64_t>(k, n);
  std::partial_sort(perm.begin(), perm.begin() + keep, perm.end(), less);
  std::vector<int64_t> kb(std::max<uint64_t>(keep, 1) * d.n_keys);
  for (uint64_t r = 0; r < keep; ++r)
    for (uint32_t q = 0; q < d.n_keys; ++q) kb[r * d.n_keys + q] = kbuf[perm[r] * d.n_keys + q];
  kbuf.swap(kb);
  for (uint32_t a = 0; a < d.n_aggs; ++a) {
    const uint64_t w = agg_width(a);
    std::vector<uint8_t> ab(std::max<uint64_t>(keep, 1) * w);
    for (uint64_t r = 0; r < keep; ++r)
      std::memcpy(ab.data() + r * w, abuf[a].data() + perm[r] * w, w);
    abuf[a].swap(ab);
    aptr[a] = abuf[a].data();
  }
  n_out = (uint32_t)keep;
  return MTL_OK;
}

// Runs the group-by over d and materialises the result columns (DuckDB's
// result types: COUNT BIGINT, integer SUM HUGEINT, AVG and float
// aggregates DOUBLE, DECIMAL SUM/MIN/MAX keeping the scale).
static int32_t emit_groups(const Query& q, mtl_group_desc& d, const GroupEmit& ge,
                           const std::vector<GroupOut>& outs, mtl_result& res, mtl_stream* s) {
  mtl_groups* g = nullptr;
  MTL_TRY(mtl_group_agg(&d, &g, s));
  struct Release {
    mtl_groups* g;
    ~Release() { mtl_groups_release(g); }
  } rel{g};
  uint64_t ng = 0;
  std::vector<int64_t> kbuf;
  std::vector<std::vector<uint8_t>> abuf(d.n_aggs);
  std::vector<void*> aptr(d.n_aggs);
  auto agg_width = [&](size_t a) -> uint64_t {
    const int32_t f = ge.agg_func[a], t = ge.agg_type[a];
    if (f == MTL_AGG_SUM && is_int_type(t)) return 16;
    return 8;
  };
  if (ge.order_agg >= 0) {
    const uint32_t k = (uint32_t)q.limit;
    kbuf.resize((uint64_t)std::max<uint32_t>(k, 1) * d.n_keys);
    for (uint32_t a = 0; a < d.n_aggs; ++a) {
      abuf[a].resize((uint64_t)std::max<uint32_t>(k, 1) * agg_width(a));
      aptr[a] = abuf[a].data();
    }
    uint32_t n = 0;
    if (k > 0 && ge.then.empty()) {
      MTL_TRY(mtl_groups_topk(g, (uint32_t)ge.order_agg, ge.odesc ? 1 : 0, k, kbuf.data(),
                              aptr.data(), &n));
    } else if (k > 0) {
      MTL_TRY(topk_then_keys(g, d, ge, k, kbuf, abuf, aptr, agg_width, n));
    }
    ng = n;
  } else {
    MTL_TRY(mtl_groups_count(g, &ng));
    kbuf.resize(std::max<uint64_t>(ng, 1) * d.n_keys);
    for (uint32_t a = 0; a < d.n_aggs; ++a) {
      abuf[a].resize(std::max<uint64_t>(ng, 1) * agg_width(a));
      aptr[a] = abuf[a].data();
    }
    uint64_t got = 0;
    MTL_TRY(mtl_groups_fetch(g, kbuf.data(), aptr.data(), ng, &got));
    ng = got;
    if (ge.key_order < 0) {  // ORDER BY the keys DESC: the dense table's order reversed
      for (uint64_t i = 0, j = ng ? ng - 1 : 0; i < j; ++i, --j) {
        for (uint32_t k = 0; k < d.n_keys; ++k)
          std::swap(kbuf[i * d.n_keys + k], kbuf[j * d.n_keys + k]);
        for (uint32_t a = 0; a < d.n_aggs; ++a) {
          const uint64_t w = agg_width(a);
          std::swap_ranges(abuf[a].begin() + i * w, abuf[a].begin() + (i + 1) * w,
                           abuf[a].begin() + j * w);
        }
      }
    }
    if (q.limit >= 0 && (uint64_t)q.limit < ng) ng = (uint64_t)q.limit;
  }
  res.rows = ng;
  // valid values per group of a NULL-skipping aggregate (its hidden count)
  auto nn_count = [&](int a, uint64_t r) -> int64_t {
    uint64_t lo;
    std::memcpy(&lo, abuf[a].data() + r * agg_width(a), 8);
    return (int64_t)lo;
  };
  for (auto& o : outs) {
    mtl_result::Column col;
    col.name = o.name;
    if (o.key >= 0) {
      const CatCol* c = ge.keys[o.key];
      col.type = c->type;
      col.scale = c->scale;
      col.dict = c->dict;
      const int w = type_width(c->type);
      col.data.resize(ng * w);
      const int64_t null_key =
          (size_t)o.key < ge.key_null.size() ? ge.key_null[o.key] : INT64_MIN;
      if (null_key != INT64_MIN) col.valid.assign(ng, 1);
      for (uint64_t r = 0; r < ng; ++r) {
        int64_t v = kbuf[r * d.n_keys + o.key];
        if (null_key != INT64_MIN && v == null_key) {
          col.valid[r] = 0;
          v = 0;
        }
        std::memcpy(col.data.data() + r * w, &v, w);  // little-endian narrowing
      }
    } else if (o.mode == 1) {  // COUNT(col): the engine's SUM of 0/1
      col.type = MTL_BIGINT;
      col.data.resize(ng * 8);
      for (uint64_t r = 0; r < ng; ++r) {
        const int64_t v = nn_count(o.agg, r);
        std::memcpy(col.data.data() + r * 8, &v, 8);
      }
    } else if (o.mode == 2) {  // AVG skipping NULLs: SUM / COUNT(col)
      col.type = MTL_DOUBLE;
      col.data.resize(ng * 8);
      col.valid.assign(ng, 1);
      const int32_t t = ge.agg_type[o.agg];
      const uint64_t w = agg_width(o.agg);
      for (uint64_t r = 0; r < ng; ++r) {
        const int64_t cnt = nn_count(o.nn, r);
        double v = 0;
        if (cnt == 0) {
          col.valid[r] = 0;
        } else if (w == 16) {
          // as the group engine's integer / DECIMAL AVG: the 128-bit sum over
          // count x 10^scale in long double
          uint64_t lo;
          int64_t hi;
          std::memcpy(&lo, abuf[o.agg].data() + r * 16, 8);
          std::memcpy(&hi, abuf[o.agg].data() + r * 16 + 8, 8);
          const __int128 sum = (__int128)(((unsigned __int128)(uint64_t)hi << 64) | lo);
          long double div = (long double)cnt;
          if (t == MTL_DECIMAL)
            for (int k = 0; k < ge.agg_scale[o.agg]; ++k) div *= 10.0L;
          v = (double)((long double)sum / div);
        } else {
          double sum;
          std::memcpy(&sum, abuf[o.agg].data() + r * 8, 8);
          v = sum / (double)cnt;
        }
        std::memcpy(col.data.data() + r * 8, &v, 8);
      }
    } else {
      const int32_t f = ge.agg_func[o.agg], t = ge.agg_type[o.agg];
      const uint64_t w = agg_width(o.agg);
      agg_out_type(f, t, ge.agg_scale[o.agg], col.type, col.scale);
      col.data.assign(abuf[o.agg].begin(), abuf[o.agg].begin() + ng * w);
      if (o.nn >= 0) {  // NULL when the group has no valid value
        col.valid.assign(ng, 1);
        for (uint64_t r = 0; r < ng; ++r) col.valid[r] = nn_count(o.nn, r) > 0;
      }
    }
    res.cols.push_back(std::move(col));
  }
  return MTL_OK;
}

// ---------------------------------------------------------------------------
// Shape 2: single-table GROUP BY
// ---------------------------------------------------------------------------
static int32_t run_group(const Query& q, Bound& b, mtl_result& res, mtl_stream* s, bool dry) {
  const CatTable& T = *b.tabs[0];
  std::vector<std::vector<mtl_pred>> tp;
  std::vector<uint32_t> tor;
  MTL_TRY(bind_preds(q, b, tp, tor));
  const std::vector<mtl_pred>& preds = tp[0];
  mtl_group_desc d{};
  d.n_rows = T.rows;
  std::vector<const CatCol*> keys;
  for (auto& g : q.group) {
    int src;
    const CatCol* c;
    if (!b.resolve(g, src, c)) return set_error(MTL_UNSUPPORTED, "%s", b.err.c_str());
    if (!is_int_type(c->type)) return set_error(MTL_UNSUPPORTED, "GROUP BY a non-integer column");
    if (c->valid) return set_error(MTL_UNSUPPORTED, "GROUP BY a column with NULLs (join shape)");
    keys.push_back(c);
  }
  if (keys.empty() || keys.size() > MTL_GROUP_MAX_KEYS)
    return set_error(MTL_UNSUPPORTED, "GROUP BY needs 1..%d columns", MTL_GROUP_MAX_KEYS);
  d.n_keys = (uint32_t)keys.size();
  for (size_t i = 0; i < keys.size(); ++i) {
    d.key_col[i] = keys[i]->dev;
    d.key_type[i] = keys[i]->type;
    d.key_min[i] = keys[i]->min;
    d.key_max[i] = keys[i]->max;
  }
  // select items: group columns or aggregates over a column
  using Out = GroupOut;
  std::vector<Out> outs;
  std::vector<int32_t> agg_func, agg_type;
  std::vector<int32_t> agg_scale;
  for (size_t i = 0; i < q.items.size(); ++i) {
    const Item& it = q.items[i];
    Out o;
    o.name = item_name(it, i);
    if (it.e->k == Expr::COL) {
      int src;
      const CatCol* c;
      if (!b.resolve(it.e->col, src, c)) return set_error(MTL_UNSUPPORTED, "%s", b.err.c_str());
      for (size_t k = 0; k < keys.size(); ++k)
        if (keys[k] == c) o.key = (int)k;
      if (o.key < 0) return set_error(MTL_UNSUPPORTED, "%s is neither grouped nor aggregated",
                                      c->name.c_str());
    } else if (it.e->k == Expr::AGG) {
      if (d.n_aggs >= MTL_GROUP_MAX_AGGS) return set_error(MTL_UNSUPPORTED, "too many aggregates");
      mtl_agg_spec& a = d.aggs[d.n_aggs];
      const std::string& f = it.e->fn;
      a.func = f == "sum" ? MTL_AGG_SUM : f == "count" ? MTL_AGG_COUNT : f == "min" ? MTL_AGG_MIN
               : f == "max" ? MTL_AGG_MAX : MTL_AGG_AVG;
      int32_t scale = 0;
      if (it.e->a) {
        if (it.e->a->k != Expr::COL)
          return set_error(MTL_UNSUPPORTED, "aggregate over an expression");
        int src;
        const CatCol* c;
        if (!b.resolve(it.e->a->col, src, c)) return set_error(MTL_UNSUPPORTED, "%s", b.err.c_str());
        if (c->type == MTL_DICT8 || c->type == MTL_DICT32)
          return set_error(MTL_UNSUPPORTED, "aggregate over a dictionary column");
        if (c->valid) return set_error(MTL_UNSUPPORTED, "aggregate over a column with NULLs (join shape)");
        a.values = c->dev;
        a.type = c->type;
        scale = c->scale;
        a.scale = c->type == MTL_DECIMAL ? c->scale : 0;
      } else if (a.func != MTL_AGG_COUNT) {
        return set_error(MTL_UNSUPPORTED, "%s(*)", f.c_str());
      }
      agg_func.push_back(a.func);
      agg_type.push_back(a.type);
      agg_scale.push_back(scale);
      o.agg = (int)d.n_aggs++;
    } else {
      return set_error(MTL_UNSUPPORTED, "select item %zu of a GROUP BY", i);
    }
    outs.push_back(o);
  }
  // ORDER BY one aggregate (by alias or the same call) [+ LIMIT]
  int order_agg = -1;
  bool odesc = true;
  if (q.order.size() > 1) return set_error(MTL_UNSUPPORTED, "ORDER BY more than one term");
  if (q.order.size() == 1) {
    const ExprP& e = q.order[0].e;
    for (size_t i = 0; i < q.items.size(); ++i) {
      const Item& it = q.items[i];
      if (outs[i].agg < 0) continue;
      const bool by_alias = e->k == Expr::COL && e->col.rel.empty() && e->col.name == outs[i].name;
      const bool same_call = e->k == Expr::AGG && e->fn == it.e->fn &&
                             ((!e->a && !it.e->a) ||
                              (e->a && it.e->a && e->a->k == Expr::COL &&
                               it.e->a->k == Expr::COL && e->a->col.name == it.e->a->col.name));
      if (by_alias || same_call) order_agg = outs[i].agg;
    }
    if (order_agg < 0) return set_error(MTL_UNSUPPORTED, "ORDER BY must name an aggregate");
    odesc = q.order[0].desc;
    if (q.limit < 0) return set_error(MTL_UNSUPPORTED, "ORDER BY an aggregate needs LIMIT");
    if (q.limit > 0xFFFFFFFFll)
      return set_error(MTL_UNSUPPORTED, "LIMIT %lld exceeds 2^32-1", (long long)q.limit);
  }
  res.shape = "group";
  res.scratch = (preds.empty() ? 0 : T.rows * 4) + group_scratch_bytes(d);
  if (dry) return MTL_OK;
  DevBuf sel, cnt;
  if (!preds.empty()) {
    MTL_TRY(sel.alloc(std::max<uint64_t>(T.rows, 1) * 4));
    MTL_TRY(cnt.alloc(16));
    MTL_TRY(mtl_filter_or(preds.data(), (uint32_t)preds.size(), tor[0], T.rows,
                          (uint32_t*)sel.p, (uint32_t*)cnt.p, s));
    uint32_t n_sel = 0;
    MTL_TRY(mtl_memcpy_d2h(&n_sel, cnt.p, 4, s));
    if (n_sel > T.rows) return set_error(MTL_ERR_PROCESSING, "device scan failed");
    d.sel = (const uint32_t*)sel.p;
    d.sel_count = (const uint32_t*)cnt.p;
    d.sel_capacity = T.rows;
  }
  GroupEmit ge{keys, agg_func, agg_type, agg_scale, order_agg, odesc, 0};
  return emit_groups(q, d, ge, outs, res, s);
}

// ---------------------------------------------------------------------------
// Shape 3: star join -> SUM -> top-K
// ---------------------------------------------------------------------------
// Postfix expression over the columns of table `fact` (fact < 0: of any
// table, each column's table recorded in *srcs); DECIMAL operands of + and -
// must share a scale (constants are rescaled), * adds scales.
static bool to_postfix(const ExprP& e, Bound& b, int fact, mtl_expr& x, int& scale,
                       std::vector<const CatCol*>& cols, std::string& err,
                       std::vector<int>* srcs = nullptr) {
  if (x.n_ops >= MTL_EXPR_MAX_OPS) return err = "expression too long", false;
  switch (e->k) {
    case Expr::COL: {
      int src;
      const CatCol* c;
      if (!b.resolve(e->col, src, c)) return err = b.err, false;
      if (fact >= 0 && src != fact) return err = "SUM over a column of another table", false;
      if (!is_int_type(c->type) || c->type == MTL_DICT8 || c->type == MTL_DICT32 ||
          c->type == MTL_DATE || c->type == MTL_TIMESTAMP)
        return err = "SUM over a non-numeric integer column", false;
      int idx = -1;
      for (size_t i = 0; i < cols.size(); ++i)
        if (cols[i] == c && (!srcs || (*srcs)[i] == src)) idx = (int)i;
      if (idx < 0) {
        if (cols.size() >= MTL_EXPR_MAX_COLS) return err = "too many columns in SUM", false;
        idx = (int)cols.size();
        cols.push_back(c);
        if (srcs) srcs->push_back(src);
        x.col[idx] = c->dev;
        x.col_type[idx] = c->type;
        x.n_cols = (uint32_t)cols.size();
      }
      x.op[x.n_ops] = MTL_EXPR_COL;
      x.arg[x.n_ops++] = idx;
      scale = c->type == MTL_DECIMAL ? c->scale : 0;
      return true;
    }
    case Expr::LIT:
      if (e->lit.k != Lit::INT) return err = "non-integer constant in SUM", false;
      x.op[x.n_ops] = MTL_EXPR_CONST;
      x.arg[x.n_ops++] = e->lit.i;
      scale = -1;  // a constant adopts its partner's scale
      return true;
    case Expr::ADD:
    case Expr::SUB:
    case Expr::MUL: {
      const uint32_t start = x.n_ops;
      int sa, sb;
      if (!to_postfix(e->a, b, fact, x, sa, cols, err, srcs)) return false;
      const uint32_t mid = x.n_ops;
      if (!to_postfix(e->b, b, fact, x, sb, cols, err, srcs)) return false;
      if (e->k == Expr::MUL) {
        scale = (sa < 0 ? 0 : sa) + (sb < 0 ? 0 : sb);
      } else {
        // rescale a bare constant operand to the other side's scale
        auto rescale = [&](uint32_t at, int to) {
          if (x.op[at] != MTL_EXPR_CONST || x.n_ops == 0) return true;
          int64_t v;
          if (!pow10_scale(x.arg[at], to, v)) return false;
          x.arg[at] = v;
          return true;
        };
        if (sa < 0 && sb >= 0 && !rescale(start, sb)) return err = "constant overflow", false;
        if (sb < 0 && sa >= 0 && !rescale(mid, sa)) return err = "constant overflow", false;
        if (sa >= 0 && sb >= 0 && sa != sb)
          return err = "adding DECIMAL values of different scales", false;
        scale = sa >= 0 ? sa : sb;
      }
      if (x.n_ops >= MTL_EXPR_MAX_OPS) return err = "expression too long", false;
      x.op[x.n_ops++] = e->k == Expr::ADD ? MTL_EXPR_ADD : e->k == Expr::SUB ? MTL_EXPR_SUB
                                                                              : MTL_EXPR_MUL;
      return true;
    }
    default:
      return err = "unsupported SUM expression", false;
  }
}

static int32_t run_star(const Query& q, Bound& b, mtl_result& res, mtl_stream* s, bool dry) {
  const int nt = (int)b.tabs.size();
  if (any_or(q)) return set_error(MTL_UNSUPPORTED, "OR in a star query");
  for (auto& pr : q.preds)
    if (pr.nulls) return set_error(MTL_UNSUPPORTED, "IS NULL in a star query");
  for (auto* t : b.tabs)
    for (auto& c : t->cols)
      if (c.valid) return set_error(MTL_UNSUPPORTED, "star query over a table with NULLs");
  // join edges -> (child source, fk column) -> (parent source, unique key)
  struct Edge {
    int child, parent;
    const CatCol *fk, *key;
  };
  std::vector<Edge> edges;
  for (auto& j : q.joins) {
    int sa, sb;
    const CatCol *ca, *cb;
    if (!b.resolve(j.l, sa, ca) || !b.resolve(j.r, sb, cb))
      return set_error(MTL_UNSUPPORTED, "%s", b.err.c_str());
    if (sa == sb) return set_error(MTL_UNSUPPORTED, "join within one table");
    if (cb->unique && !ca->unique) edges.push_back({sa, sb, ca, cb});
    else if (ca->unique && !cb->unique) edges.push_back({sb, sa, cb, ca});
    else if (ca->unique && cb->unique)
      return set_error(MTL_UNSUPPORTED, "one-to-one join");
    else return set_error(MTL_UNSUPPORTED, "join without a unique key side");
  }
  if ((int)edges.size() != nt - 1) return set_error(MTL_UNSUPPORTED, "joins do not chain the tables");
  // chain fact -> parent [-> dimension]
  int fact = -1;
  for (int i = 0; i < nt; ++i) {
    bool is_parent = false;
    for (auto& e : edges) is_parent |= e.parent == i;
    if (!is_parent) {
      if (fact >= 0) return set_error(MTL_UNSUPPORTED, "more than one fact table");
      fact = i;
    }
  }
  if (fact < 0) return set_error(MTL_UNSUPPORTED, "cyclic joins");
  const Edge* fe = nullptr;
  const Edge* pe = nullptr;
  for (auto& e : edges)
    if (e.child == fact) fe = &e;
  if (!fe) return set_error(MTL_UNSUPPORTED, "fact table not joined");
  const int parent = fe->parent;
  for (auto& e : edges)
    if (&e != fe) {
      if (e.child != parent) return set_error(MTL_UNSUPPORTED, "join graph is not a chain");
      pe = &e;
    }
  const int dim = pe ? pe->parent : -1;
  // group by: parent columns including its key
  const CatCol* pkey = fe->key;
  bool has_key = false;
  std::vector<const CatCol*> gcols;
  for (auto& g : q.group) {
    int src;
    const CatCol* c;
    if (!b.resolve(g, src, c)) return set_error(MTL_UNSUPPORTED, "%s", b.err.c_str());
    if (src == fact && c == fe->fk) c = pkey;  // l_orderkey = o_orderkey
    else if (src != parent) return set_error(MTL_UNSUPPORTED, "GROUP BY a non-parent column");
    has_key |= c == pkey;
    gcols.push_back(c);
  }
  if (!has_key) return set_error(MTL_UNSUPPORTED, "GROUP BY must include the parent key");
  // select items: group columns and exactly one SUM
  mtl_star_desc d{};
  int sum_item = -1, sum_scale = 0;
  struct Out {
    std::string name;
    const CatCol* c = nullptr;  // parent column; null = the sum
  };
  std::vector<Out> outs;
  for (size_t i = 0; i < q.items.size(); ++i) {
    const Item& it = q.items[i];
    Out o;
    o.name = item_name(it, i);
    if (it.e->k == Expr::COL) {
      int src;
      const CatCol* c;
      if (!b.resolve(it.e->col, src, c)) return set_error(MTL_UNSUPPORTED, "%s", b.err.c_str());
      if (src == fact && c == fe->fk) c = pkey;
      bool grouped = false;
      for (auto* g : gcols) grouped |= g == c;
      if (!grouped) return set_error(MTL_UNSUPPORTED, "%s is not grouped", c->name.c_str());
      o.c = c;
    } else if (it.e->k == Expr::AGG && it.e->fn == "sum" && sum_item < 0) {
      std::vector<const CatCol*> cols;
      std::string err;
      d.agg = mtl_expr{};
      if (!to_postfix(it.e->a, b, fact, d.agg, sum_scale, cols, err))
        return set_error(MTL_UNSUPPORTED, "%s", err.c_str());
      if (sum_scale < 0) sum_scale = 0;
      sum_item = (int)i;
    } else {
      return set_error(MTL_UNSUPPORTED, "select item %zu of the star query", i);
    }
    outs.push_back(o);
  }
  if (sum_item < 0) return set_error(MTL_UNSUPPORTED, "no SUM in the star query");
  // ORDER BY sum DESC, then parent columns
  if (q.order.empty() || q.limit < 0) return set_error(MTL_UNSUPPORTED, "star query needs ORDER BY ... LIMIT");
  {
    const ExprP& e = q.order[0].e;
    const Item& si = q.items[sum_item];
    const bool by_alias = e->k == Expr::COL && e->col.rel.empty() && e->col.name == outs[sum_item].name;
    const bool same = e->k == Expr::AGG && e->fn == "sum";
    if (!(by_alias || same) || !q.order[0].desc)
      return set_error(MTL_UNSUPPORTED, "the star query orders by its SUM descending first");
    (void)si;
  }
  for (size_t i = 1; i < q.order.size(); ++i) {
    const ExprP& e = q.order[i].e;
    if (e->k != Expr::COL) return set_error(MTL_UNSUPPORTED, "ORDER BY an expression");
    int src;
    const CatCol* c = nullptr;
    for (auto& o : outs)
      if (e->col.rel.empty() && o.c && o.name == e->col.name) c = o.c;
    if (!c) {
      if (!b.resolve(e->col, src, c)) return set_error(MTL_UNSUPPORTED, "%s", b.err.c_str());
      if (src == fact && c == fe->fk) c = pkey;
      else if (src != parent) return set_error(MTL_UNSUPPORTED, "ORDER BY a non-parent column");
    }
    if (c == pkey && !q.order[i].desc) break;  // the operator's final tie-break
    if (d.n_ties >= MTL_STAR_MAX_TIES) return set_error(MTL_UNSUPPORTED, "too many ORDER BY terms");
    d.tie_col[d.n_ties] = c->dev;
    d.tie_type[d.n_ties] = c->type;
    d.tie_desc[d.n_ties] = q.order[i].desc ? 1 : 0;
    ++d.n_ties;
  }
  if (q.limit > 0xFFFFFFFFll) return set_error(MTL_UNSUPPORTED, "LIMIT %lld", (long long)q.limit);
  d.k = (uint32_t)q.limit;  // the operator rejects K it cannot serve
  // filters by table
  auto add_pred = [&](int src, mtl_pred* arr, uint32_t& n, const Pred& pr,
                      const CatCol* c) -> int32_t {
    if (n >= MTL_MAX_PREDS) return set_error(MTL_UNSUPPORTED, "too many predicates");
    mtl_pred mp{};
    mp.column = c->dev;
    mp.type = c->type;
    mp.op = pr.op;
    std::string err;
    if (!literal_bits(pr.lit, *c, mp.op, mp.scalar, err))
      return set_error(MTL_UNSUPPORTED, "%s", err.c_str());
    arr[n++] = mp;
    (void)src;
    return MTL_OK;
  };
  for (auto& pr : q.preds) {
    int src;
    const CatCol* c;
    if (!b.resolve(pr.col, src, c)) return set_error(MTL_UNSUPPORTED, "%s", b.err.c_str());
    if (src == fact) MTL_TRY(add_pred(src, d.f_preds, d.f_n_preds, pr, c));
    else if (src == parent) MTL_TRY(add_pred(src, d.p_preds, d.p_n_preds, pr, c));
    else MTL_TRY(add_pred(src, d.d_preds, d.d_n_preds, pr, c));
  }
  const CatTable& F = *b.tabs[fact];
  const CatTable& P = *b.tabs[parent];
  d.f_rows = F.rows;
  d.f_fk = fe->fk->dev;
  d.f_fk_type = fe->fk->type;
  d.p_rows = P.rows;
  d.p_key = pkey->dev;
  d.p_key_type = pkey->type;
  d.p_key_min = pkey->min;
  d.p_key_max = pkey->max;
  if (dim >= 0) {
    const CatTable& D = *b.tabs[dim];
    d.d_rows = D.rows;
    d.d_key = pe->key->dev;
    d.d_key_type = pe->key->type;
    d.d_key_min = pe->key->min;
    d.d_key_max = pe->key->max;
    d.p_fk = pe->fk->dev;
    d.p_fk_type = pe->fk->type;
  }
  res.shape = dim >= 0 ? "star3" : "star2";
  {
    mtl_query sq{};
    sq.kind = MTL_Q_STAR;
    sq.star = d;
    res.scratch = query_scratch(sq);
  }
  if (dry) return MTL_OK;
  std::vector<mtl_star_row> rows(std::max<uint32_t>(d.k, 1));
  uint32_t n = 0;
  if (d.k) MTL_TRY(mtl_star_topk(&d, rows.data(), &n, s));
  res.rows = n;
  // gather the group columns of the K parent rows on the device
  DevBuf prow, out;
  MTL_TRY(prow.alloc((uint64_t)std::max<uint32_t>(n, 1) * 4));
  MTL_TRY(out.alloc((uint64_t)std::max<uint32_t>(n, 1) * 8));
  std::vector<uint32_t> ids(n);
  for (uint32_t i = 0; i < n; ++i) ids[i] = rows[i].p_row;
  if (n) MTL_TRY(mtl_memcpy_h2d(prow.p, ids.data(), (uint64_t)n * 4, s));
  for (auto& o : outs) {
    mtl_result::Column col;
    col.name = o.name;
    if (o.c) {
      col.type = o.c->type;
      col.scale = o.c->scale;
      col.dict = o.c->dict;
      const uint32_t w = (uint32_t)type_width(o.c->type);
      if (n) MTL_TRY(mtl_gather(o.c->dev, w, (const uint32_t*)prow.p, nullptr, n, out.p, s));
      MTL_TRY(download(col.data, out.p, (uint64_t)n * w, s));
    } else {
      col.type = d.agg.is_double ? MTL_DOUBLE : MTL_DECIMAL;
      col.scale = sum_scale;
      col.data.resize((uint64_t)n * 8);
      for (uint32_t i = 0; i < n; ++i) std::memcpy(col.data.data() + i * 8, &rows[i].sum_i64, 8);
    }
    res.cols.push_back(std::move(col));
  }
  return MTL_OK;
}

// ---------------------------------------------------------------------------
// Shape 4: tables joined along equi-join edges that form a tree (each with
// its own filters), then a projection of columns, a GROUP BY over columns of
// any table, or aggregates without GROUP BY (one table or several).
//   filters   : mtl_filter per table
//   joins     : the relation starts at the table with the most filtered
//               rows; each edge to a new table builds a join table over the
//               side whose key is unique (the new table's filtered rows, else
//               the relation so far), probes it with the other side, and
//               gathers the row ids every relation column carries
//               (mtl_join_build / mtl_join_probe / mtl_gather)
//   outputs   : mtl_gather of each column through its table's row ids;
//               ORDER BY one column with mtl_sort; aggregates by
//               mtl_group_agg over the gathered key and value columns
// Reference: the generic GPU plan (executor/mod.rs:257-386, join order and
// strategies join_plan.rs:58-298) and ungrouped aggregates
// (aggregate.rs:424-455).  Joins where neither side's key is unique (the
// reference overwrites duplicate build keys, kernels.metal:202-206) return
// MTL_UNSUPPORTED.
// ---------------------------------------------------------------------------
// Multi-column join keys packed into one int64: sum over k of (key_k -
// min_k) * mult_k (mixed radix over each key pair's combined range), for rows
// sel[i] (i without a selection), i < n.
struct PackKeys {
  const void* col[MTL_GROUP_MAX_KEYS];
  int32_t type[MTL_GROUP_MAX_KEYS];
  int64_t min[MTL_GROUP_MAX_KEYS];
  uint64_t mult[MTL_GROUP_MAX_KEYS];
  uint32_t n_keys;
};

__global__ void k_pack_keys(PackKeys pk, const uint32_t* sel, uint64_t n, int64_t* out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = sel ? sel[i] : (uint32_t)i;
    uint64_t v = 0;
    for (uint32_t k = 0; k < pk.n_keys; ++k)
      v += (uint64_t)(dev::load_key(pk.col[k], pk.type[k], r) - pk.min[k]) * pk.mult[k];
    out[i] = (int64_t)v;
  }
}

static int32_t pack_keys(const PackKeys& pk, const uint32_t* sel, uint64_t n, DevBuf& out,
                         mtl_stream* s) {
  MTL_TRY(out.alloc(std::max<uint64_t>(n, 1) * 8));
  if (!n) return MTL_OK;
  const uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(n, 256), 4096);
  hipLaunchKernelGGL(k_pack_keys, dim3(grid), dim3(256), 0, s->stream, pk, sel, n,
                     (int64_t*)out.p);
  MTL_HIP(hipGetLastError());
  return MTL_OK;
}

// NULL handling over gathered columns (one byte of validity per row).
// acc[i] &= v[i]: the validity of an expression is that of all its columns.
__global__ void k_valid_and(uint8_t* acc, const uint8_t* v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    acc[i] &= v[i];
}
// col[i] = fill where valid[i] == 0 (width 4 or 8): the identity of an
// aggregate (0 for SUM, the type's maximum for MIN, ...) replaces NULLs.
__global__ void k_null_fill(void* col, uint32_t width, const uint8_t* valid, uint64_t n,
                            uint64_t fill) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (valid[i]) continue;
    if (width == 8) ((uint64_t*)col)[i] = fill;
    else ((uint32_t*)col)[i] = (uint32_t)fill;
  }
}
// out[i] = valid[i] (int32): COUNT(col) as the SUM of this column.
__global__ void k_valid_i32(const uint8_t* valid, uint64_t n, int32_t* out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = valid[i];
}

static uint32_t grid_for(uint64_t n) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, 256), 4096));
}

struct JoinRel {
  std::vector<std::unique_ptr<DevBuf>> ids;  // per source: row ids (nullptr: not joined)
  uint64_t m = 0;                            // rows of the relation
};

// Build a join table over (bkeys, btype) through bsel (brows rows, keys in
// [bmin, bmax]) and probe it with (pkeys, ptype) through psel (prows rows;
// NULL: rows 0..prows-1): (probe row, build row) pairs, in probe order when
// `ordered` (else a large build takes the radix-partitioned join, whose
// pairs come in partition order).  MTL_UNSUPPORTED when a build key repeats.
static int32_t join_pairs(const void* bkeys, int32_t btype, const uint32_t* bsel, uint64_t brows,
                          int64_t bmin, int64_t bmax, const void* pkeys, int32_t ptype,
                          const uint32_t* psel, uint64_t prows, DevBuf& out_p, DevBuf& out_b,
                          uint64_t& n_pairs, mtl_stream* s, bool ordered = true) {
  n_pairs = 0;
  MTL_TRY(out_p.alloc(std::max<uint64_t>(prows, 1) * 4));
  MTL_TRY(out_b.alloc(std::max<uint64_t>(prows, 1) * 4));
  if (brows == 0 || prows == 0) return MTL_OK;
  // the dense-vs-hash rule of join_plan.rs:29-56: build >= 1024 rows, key
  // span <= 2^24 and <= max(16 x build, 4096)
  const uint64_t span = (uint64_t)bmax - (uint64_t)bmin + 1;
  const bool dense = bmax >= bmin && brows >= 1024 && span <= (1ull << 24) &&
                     span <= std::max<uint64_t>(16 * brows, 4096);
  // radix partitions beat the global hash table from a few thousand build
  // rows on (no random HBM lookups), up to its 2^24-row limit
  int32_t strategy = dense ? MTL_JOIN_DENSE
                     : (!ordered && brows >= 4096 && brows <= (1ull << 24)) ? MTL_JOIN_RADIX
                                                                           : MTL_JOIN_HASH;
  const uint64_t slots = dense ? span : brows;
  DevBuf table, flag, cnt;
  MTL_TRY(flag.alloc(16));
  MTL_TRY(cnt.alloc(16));
  const uint32_t zero = 0, br = (uint32_t)brows;
  MTL_TRY(mtl_memcpy_h2d(cnt.p, &br, 4, s));  // the build selection's count
  uint32_t err = 0;
  uint64_t tb = 0;
  for (;;) {
    tb = mtl_join_table_bytes(strategy, brows, dense ? span : 0);
    MTL_TRY(table.alloc(tb));
    MTL_TRY(mtl_memcpy_h2d(flag.p, &zero, 4, s));
    MTL_TRY(mtl_join_build(strategy, bkeys, btype, bsel, bsel ? (const uint32_t*)cnt.p : nullptr,
                           brows, bmin, slots, table.p, tb, (uint32_t*)flag.p, s));
    MTL_TRY(mtl_memcpy_d2h(&err, flag.p, 4, s));
    if (!(err & 16u) || strategy != MTL_JOIN_RADIX) break;
    strategy = MTL_JOIN_HASH;  // a partition over the LDS table: the global table
  }
  if (err & 2u) return set_error(MTL_UNSUPPORTED, "join with duplicate keys on the build side");
  if (err & 8u) return set_error(MTL_ERR_PROCESSING, "join key outside the column's stated range");
  DevBuf pcnt;
  MTL_TRY(pcnt.alloc(16));
  uint32_t pr = (uint32_t)prows;
  MTL_TRY(mtl_memcpy_h2d(pcnt.p, &pr, 4, s));
  MTL_TRY(mtl_join_probe(strategy, table.p, tb, bmin, slots, pkeys, ptype, psel,
                         psel ? (const uint32_t*)pcnt.p : nullptr, prows, (uint32_t*)out_p.p,
                         (uint32_t*)out_b.p, (uint32_t*)cnt.p, s));
  uint32_t n = 0;
  MTL_TRY(mtl_memcpy_d2h(&n, cnt.p, 4, s));
  if (n > prows) return set_error(MTL_ERR_PROCESSING, "device join failed");
  n_pairs = n;
  return MTL_OK;
}

// Many-to-many join: every (probe, build) pair of equal keys, in probe order
// and, within a probe row, in build-selection order.  The build side is
// sorted by key into runs; join_pairs over the runs' unique keys gives
// (probe, run) pairs, which expand to one pair per build row of the run.
static int32_t join_pairs_multi(const void* bkeys, int32_t btype, const uint32_t* bsel,
                                uint64_t brows, int64_t bmin, int64_t bmax, const void* pkeys,
                                int32_t ptype, const uint32_t* psel, uint64_t prows,
                                DevBuf& out_p, DevBuf& out_b, uint64_t& n_pairs, mtl_stream* s,
                                bool ordered = true) {
  n_pairs = 0;
  if (brows == 0 || prows == 0) {
    MTL_TRY(out_p.alloc(4));
    MTL_TRY(out_b.alloc(4));
    return MTL_OK;
  }
  if (brows > 0xFFFFFFFEull) return set_error(MTL_UNSUPPORTED, "join build over 2^32 rows");
  // build rows in key order (stable), and their keys
  DevBuf srow, skeys, cnt;
  MTL_TRY(srow.alloc(brows * 4));
  MTL_TRY(cnt.alloc(16));
  const uint32_t br = (uint32_t)brows;
  MTL_TRY(mtl_memcpy_h2d(cnt.p, &br, 4, s));
  uint32_t ns = 0;
  MTL_TRY(mtl_sort(bkeys, btype, bsel, bsel ? (const uint32_t*)cnt.p : nullptr, brows, 0,
                   (uint32_t*)srow.p, &ns, s));
  if (ns != brows) return set_error(MTL_ERR_PROCESSING, "join build sort returned %u rows", ns);
  const uint32_t w = (uint32_t)type_width(btype);
  MTL_TRY(skeys.alloc(brows * w));
  MTL_TRY(mtl_gather(bkeys, w, (const uint32_t*)srow.p, nullptr, brows, skeys.p, s));
  // runs of equal keys
  const uint64_t tiles = (brows + 4095) / 4096 + 1;
  DevBuf tb, rstart, rkey;
  MTL_TRY(tb.alloc(tiles * 4));
  MTL_TRY(rstart.alloc((brows + 1) * 4));
  MTL_TRY(rkey.alloc(brows * 8));
  MTL_TRY(join_runs(skeys.p, btype, brows, (uint32_t*)tb.p, (uint32_t*)rstart.p,
                    (int64_t*)rkey.p, (uint32_t*)cnt.p, s->stream));
  uint32_t runs = 0;
  MTL_TRY(mtl_memcpy_d2h(&runs, cnt.p, 4, s));
  if (runs == 0 || runs > brows) return set_error(MTL_ERR_PROCESSING, "join runs: %u", runs);
  // (probe, run) pairs over the unique run keys
  DevBuf pp, pr;
  uint64_t np = 0;
  MTL_TRY(join_pairs(rkey.p, MTL_BIGINT, nullptr, runs, bmin, bmax, pkeys, ptype, psel, prows, pp,
                     pr, np, s, ordered));
  // expand each pair to its run's rows
  const uint64_t ptiles = (np + 4095) / 4096 + 1;
  DevBuf ptb, total;
  MTL_TRY(ptb.alloc(ptiles * 4));
  MTL_TRY(total.alloc(16));
  const uint64_t zero[2] = {0, 0};
  MTL_TRY(mtl_memcpy_h2d(total.p, zero, 16, s));
  MTL_TRY(join_expand_count((const uint32_t*)pr.p, np, (const uint32_t*)rstart.p,
                            (uint32_t*)ptb.p, (unsigned long long*)total.p, s->stream));
  uint64_t t = 0;
  MTL_TRY(mtl_memcpy_d2h(&t, total.p, 8, s));
  if (t > 0xFFFFFFFFull)
    return set_error(MTL_UNSUPPORTED, "join result of %llu rows exceeds 2^32-1",
                     (unsigned long long)t);
  MTL_TRY(out_p.alloc(std::max<uint64_t>(t, 1) * 4));
  MTL_TRY(out_b.alloc(std::max<uint64_t>(t, 1) * 4));
  DevBuf off;
  MTL_TRY(off.alloc(std::max<uint64_t>(np, 1) * 4));
  MTL_TRY(join_expand_write((const uint32_t*)pp.p, (const uint32_t*)pr.p, np,
                            (const uint32_t*)rstart.p, (const uint32_t*)srow.p, (uint32_t*)ptb.p,
                            (uint32_t*)total.p + 2, (uint32_t*)off.p, t, (uint32_t*)out_p.p,
                            (uint32_t*)out_b.p, s->stream));
  n_pairs = t;
  return MTL_OK;
}

// dst[i] = src[idx[i]] for i < n (uint32 row ids).
static int32_t gather_ids(const void* src, const void* idx, uint64_t n, DevBuf& dst,
                          mtl_stream* s) {
  MTL_TRY(dst.alloc(std::max<uint64_t>(n, 1) * 4));
  if (n) MTL_TRY(mtl_gather(src, 4, (const uint32_t*)idx, nullptr, n, dst.p, s));
  return MTL_OK;
}

static int32_t run_join(const Query& q, Bound& b, mtl_result& res, mtl_stream* s, bool dry) {
  const size_t ns = b.tabs.size();
  if (ns > 4) return set_error(MTL_UNSUPPORTED, "more than four joined tables");
  std::vector<std::vector<mtl_pred>> preds;
  std::vector<uint32_t> pred_or;
  MTL_TRY(bind_preds(q, b, preds, pred_or));
  // one edge per joined pair of tables; several equalities between the same
  // pair are a multi-column key
  struct Edge {
    int a, c;
    const CatCol *ca, *cc;              // the first key pair
    std::vector<const CatCol*> xa, xc;  // every key pair (a side, c side)
    bool used = false;
  };
  std::vector<Edge> edges;
  for (auto& j : q.joins) {
    Edge e;
    if (!b.resolve(j.l, e.a, e.ca) || !b.resolve(j.r, e.c, e.cc))
      return set_error(MTL_UNSUPPORTED, "%s", b.err.c_str());
    if (e.a == e.c) return set_error(MTL_UNSUPPORTED, "join condition within one table");
    if (!is_int_type(e.ca->type) || !is_int_type(e.cc->type) || e.ca->type == MTL_DICT8 ||
        e.cc->type == MTL_DICT8 || e.ca->type == MTL_DICT32 || e.cc->type == MTL_DICT32)
      return set_error(MTL_UNSUPPORTED, "join on a non-integer column");
    Edge* same = nullptr;
    for (auto& x : edges)
      if ((x.a == e.a && x.c == e.c) || (x.a == e.c && x.c == e.a)) same = &x;
    if (same) {
      const bool flip = same->a != e.a;
      same->xa.push_back(flip ? e.cc : e.ca);
      same->xc.push_back(flip ? e.ca : e.cc);
      if (same->xa.size() > MTL_GROUP_MAX_KEYS)
        return set_error(MTL_UNSUPPORTED, "more than %d join keys", MTL_GROUP_MAX_KEYS);
      continue;
    }
    e.xa.push_back(e.ca);
    e.xc.push_back(e.cc);
    edges.push_back(e);
  }
  if (edges.size() != ns - 1)
    return set_error(MTL_UNSUPPORTED, "joins must connect the tables as a tree");
  for (auto& e : edges)  // NULL keys match nothing
    for (size_t k = 0; k < e.xa.size(); ++k) {
      MTL_TRY(add_not_null(preds, e.a, *e.xa[k]));
      MTL_TRY(add_not_null(preds, e.c, *e.xc[k]));
    }
  // outputs
  bool has_agg = false;
  for (auto& it : q.items) has_agg |= it.e->k == Expr::AGG;
  struct Col {
    int src;
    const CatCol* c;
  };
  std::vector<Col> proj;  // projection outputs
  std::vector<std::string> proj_name;
  std::vector<Col> gkeys;  // group keys
  std::vector<const CatCol*> gkey_cols;
  std::vector<Col> gvals;  // aggregate value columns (src -1: COUNT(*))
  struct AggExpr {
    mtl_expr x{};
    std::vector<Col> cols;  // the expression's columns and their tables
  };
  std::vector<std::unique_ptr<AggExpr>> gexpr;  // per aggregate; null: a column
  // per aggregate: the columns with NULLs its value depends on, and whether
  // the aggregate is the hidden count of their valid rows
  std::vector<std::vector<Col>> gnull;
  std::vector<bool> gcount;
  std::vector<GroupOut> gouts;
  std::vector<int32_t> agg_func, agg_type, agg_scale;
  mtl_group_desc d{};
  if (!has_agg && q.group.empty()) {
    for (size_t i = 0; i < q.items.size(); ++i) {
      const Item& it = q.items[i];
      if (it.e->k != Expr::COL)
        return set_error(MTL_UNSUPPORTED, "join select item %zu is not a column", i);
      Col c;
      if (!b.resolve(it.e->col, c.src, c.c)) return set_error(MTL_UNSUPPORTED, "%s", b.err.c_str());
      proj.push_back(c);
      proj_name.push_back(item_name(it, i));
    }
  } else {
    for (auto& g : q.group) {
      Col c;
      if (!b.resolve(g, c.src, c.c)) return set_error(MTL_UNSUPPORTED, "%s", b.err.c_str());
      if (!is_int_type(c.c->type)) return set_error(MTL_UNSUPPORTED, "GROUP BY a non-integer column");
      gkeys.push_back(c);
      gkey_cols.push_back(c.c);
    }
    if (gkeys.size() > MTL_GROUP_MAX_KEYS)
      return set_error(MTL_UNSUPPORTED, "GROUP BY more than %d columns", MTL_GROUP_MAX_KEYS);
    for (size_t i = 0; i < q.items.size(); ++i) {
      const Item& it = q.items[i];
      GroupOut o;
      o.name = item_name(it, i);
      if (it.e->k == Expr::COL) {
        Col c;
        if (!b.resolve(it.e->col, c.src, c.c)) return set_error(MTL_UNSUPPORTED, "%s", b.err.c_str());
        for (size_t k = 0; k < gkeys.size(); ++k)
          if (gkeys[k].c == c.c && gkeys[k].src == c.src) o.key = (int)k;
        if (o.key < 0)
          return set_error(MTL_UNSUPPORTED, "%s is neither grouped nor aggregated", c.c->name.c_str());
      } else if (it.e->k == Expr::AGG) {
        if (d.n_aggs >= MTL_GROUP_MAX_AGGS) return set_error(MTL_UNSUPPORTED, "too many aggregates");
        const std::string& f = it.e->fn;
        const int32_t func = f == "sum" ? MTL_AGG_SUM : f == "count" ? MTL_AGG_COUNT
                             : f == "min" ? MTL_AGG_MIN : f == "max" ? MTL_AGG_MAX : MTL_AGG_AVG;
        Col c{-1, nullptr};
        int32_t vtype = 0, vscale = 0;
        std::unique_ptr<AggExpr> ax;
        std::vector<Col> nulls;
        if (it.e->a && it.e->a->k != Expr::COL) {
          // an integer/DECIMAL expression of columns of any joined tables,
          // evaluated per joined row into an int64 column
          if (func == MTL_AGG_COUNT) return set_error(MTL_UNSUPPORTED, "COUNT of an expression");
          ax.reset(new AggExpr());
          std::vector<const CatCol*> cols;
          std::vector<int> srcs;
          std::string err;
          int sc = 0;
          if (!to_postfix(it.e->a, b, -1, ax->x, sc, cols, err, &srcs))
            return set_error(MTL_UNSUPPORTED, "%s", err.c_str());
          bool dec = false;
          for (size_t k = 0; k < cols.size(); ++k) {
            ax->cols.push_back(Col{srcs[k], cols[k]});
            dec |= cols[k]->type == MTL_DECIMAL;
          }
          if (cols.empty()) return set_error(MTL_UNSUPPORTED, "aggregate of a constant");
          vtype = dec ? MTL_DECIMAL : MTL_BIGINT;
          for (auto& xc : ax->cols)
            if (xc.c->valid) nulls.push_back(xc);
          vscale = dec ? std::max(sc, 0) : 0;
          c = Col{ax->cols[0].src, ax->cols[0].c};  // non-null: the aggregate reads values
        } else if (it.e->a) {
          if (!b.resolve(it.e->a->col, c.src, c.c)) return set_error(MTL_UNSUPPORTED, "%s", b.err.c_str());
          if (c.c->type == MTL_DICT8 || c.c->type == MTL_DICT32)
            return set_error(MTL_UNSUPPORTED, "aggregate over a dictionary column");
          vtype = c.c->type;
          vscale = c.c->type == MTL_DECIMAL ? c.c->scale : 0;
          if (c.c->valid) nulls.push_back(c);
        } else if (func != MTL_AGG_COUNT) {
          return set_error(MTL_UNSUPPORTED, "%s(*)", f.c_str());
        }
        d.aggs[d.n_aggs].func = func;
        d.aggs[d.n_aggs].type = vtype;
        d.aggs[d.n_aggs].scale = vscale;
        agg_func.push_back(func);
        agg_type.push_back(vtype);
        agg_scale.push_back(vscale);
        gvals.push_back(c);
        gexpr.push_back(std::move(ax));
        gnull.push_back(std::move(nulls));
        gcount.push_back(false);
        o.agg = (int)d.n_aggs++;
      } else {
        return set_error(MTL_UNSUPPORTED, "select item %zu of an aggregate query", i);
      }
      gouts.push_back(o);
    }
    if (gkeys.empty())
      for (auto& o : gouts)
        if (o.key >= 0) return set_error(MTL_UNSUPPORTED, "a column beside ungrouped aggregates");
    // aggregates over values with NULLs skip them: COUNT(col) becomes the
    // SUM of a 0/1 column; SUM/MIN/MAX/AVG get a count of the valid values
    // (NULL where it is 0; aggregates over the same columns share one, a
    // COUNT(col) among them included), AVG is that SUM over that count
    auto same_nulls = [](const std::vector<Col>& x, const std::vector<Col>& y) {
      if (x.size() != y.size()) return false;
      for (size_t i = 0; i < x.size(); ++i)
        if (x[i].c != y[i].c || x[i].src != y[i].src) return false;
      return true;
    };
    std::vector<int> counts;  // engine aggregates counting valid rows
    for (auto& o : gouts) {
      if (o.agg < 0 || gnull[o.agg].empty() || agg_func[o.agg] != MTL_AGG_COUNT) continue;
      const int a = o.agg;
      d.aggs[a].func = agg_func[a] = MTL_AGG_SUM;
      d.aggs[a].type = agg_type[a] = MTL_INTEGER;
      d.aggs[a].scale = agg_scale[a] = 0;
      gcount[a] = true;
      o.mode = 1;
      counts.push_back(a);
    }
    for (auto& o : gouts) {
      if (o.agg < 0 || gnull[o.agg].empty() || o.mode == 1) continue;
      const int a = o.agg;
      int h = -1;
      for (int c : counts)
        if (same_nulls(gnull[c], gnull[a])) h = c;
      if (h < 0) {
        if (d.n_aggs >= MTL_GROUP_MAX_AGGS)
          return set_error(MTL_UNSUPPORTED, "too many aggregates with NULLs");
        h = (int)d.n_aggs++;
        d.aggs[h].func = MTL_AGG_SUM;
        d.aggs[h].type = MTL_INTEGER;
        d.aggs[h].scale = 0;
        agg_func.push_back(MTL_AGG_SUM);
        agg_type.push_back(MTL_INTEGER);
        agg_scale.push_back(0);
        gvals.push_back(gnull[a][0]);
        gexpr.emplace_back();
        std::vector<Col> nl = gnull[a];
        gnull.push_back(std::move(nl));
        gcount.push_back(true);
        counts.push_back(h);
      }
      o.nn = h;
      if (agg_func[a] == MTL_AGG_AVG) {
        d.aggs[a].func = agg_func[a] = MTL_AGG_SUM;
        o.mode = 2;
      }
    }
  }
  // ORDER BY
  const bool grouped = !gkeys.empty() || has_agg;
  int order_agg = -1, key_order = 0;
  bool odesc = false;
  std::vector<std::pair<int, bool>> then_keys;
  Col okey{-1, nullptr};
  if (!grouped) {
    if (q.order.size() > 1) return set_error(MTL_UNSUPPORTED, "ORDER BY more than one column");
    if (q.order.size() == 1) {
      const ExprP& e = q.order[0].e;
      if (e->k != Expr::COL) return set_error(MTL_UNSUPPORTED, "ORDER BY an expression");
      for (size_t i = 0; i < proj.size(); ++i)
        if (e->col.rel.empty() && proj_name[i] == e->col.name) okey = proj[i];
      if (!okey.c && !b.resolve(e->col, okey.src, okey.c))
        return set_error(MTL_UNSUPPORTED, "%s", b.err.c_str());
      odesc = q.order[0].desc;
    }
  } else if (!q.order.empty()) {
    // the group keys in GROUP BY order, one direction; or one aggregate + LIMIT
    bool keys_ok = q.order.size() <= gkeys.size();
    for (size_t i = 0; keys_ok && i < q.order.size(); ++i) {
      const ExprP& e = q.order[i].e;
      keys_ok = e->k == Expr::COL && q.order[i].desc == q.order[0].desc;
      if (!keys_ok) break;
      int src;
      const CatCol* c = nullptr;
      bool hit = false;
      for (auto& o : gouts)
        if (o.key == (int)i && e->col.rel.empty() && o.name == e->col.name) hit = true;
      if (!hit && b.resolve(e->col, src, c)) hit = gkeys[i].c == c && gkeys[i].src == src;
      keys_ok = hit;
    }
    if (keys_ok) {
      key_order = q.order[0].desc ? -1 : 1;
      if (key_order < 0)
        for (auto& g : gkeys)
          if (g.c->valid)
            return set_error(MTL_UNSUPPORTED, "ORDER BY DESC over group keys with NULLs");
    } else {
      // one aggregate, then group keys as tie-breaks
      const ExprP& e = q.order[0].e;
      for (size_t i = 0; i < q.items.size(); ++i) {
        if (gouts[i].agg < 0) continue;
        const Item& it = q.items[i];
        const bool by_alias = e->k == Expr::COL && e->col.rel.empty() && e->col.name == gouts[i].name;
        const bool same_call = e->k == Expr::AGG && e->fn == it.e->fn &&
                               ((!e->a && !it.e->a) ||
                                (e->a && it.e->a && e->a->k == Expr::COL &&
                                 it.e->a->k == Expr::COL && e->a->col.name == it.e->a->col.name));
        if (by_alias || same_call) order_agg = gouts[i].agg;
      }
      if (order_agg < 0) return set_error(MTL_UNSUPPORTED, "ORDER BY of an aggregate query");
      for (auto& o : gouts)
        if (o.agg == order_agg && (o.nn >= 0 || o.mode != 0))
          return set_error(MTL_UNSUPPORTED, "ORDER BY an aggregate with NULLs");
      for (size_t i = 1; i < q.order.size(); ++i) {
        const ExprP& x = q.order[i].e;
        if (x->k != Expr::COL) return set_error(MTL_UNSUPPORTED, "ORDER BY of an aggregate query");
        int key = -1;
        for (auto& o : gouts)
          if (o.key >= 0 && x->col.rel.empty() && o.name == x->col.name) key = o.key;
        int src;
        const CatCol* c = nullptr;
        if (key < 0 && b.resolve(x->col, src, c))
          for (size_t k = 0; k < gkeys.size(); ++k)
            if (gkeys[k].c == c && gkeys[k].src == src) key = (int)k;
        if (key < 0)
          return set_error(MTL_UNSUPPORTED, "ORDER BY a column that is not a group key");
        if (q.order[i].desc && gkeys[key].c->valid)
          return set_error(MTL_UNSUPPORTED, "ORDER BY DESC over group keys with NULLs");
        then_keys.emplace_back(key, q.order[i].desc);
      }
      if (q.limit < 0) return set_error(MTL_UNSUPPORTED, "ORDER BY an aggregate needs LIMIT");
      if (q.limit > 0xFFFFFFFFll)
        return set_error(MTL_UNSUPPORTED, "LIMIT %lld exceeds 2^32-1", (long long)q.limit);
      odesc = q.order[0].desc;
    }
  }
  // group descriptor (columns filled after the joins); a key with NULLs
  // gets one more value above its range standing for NULL (NULLs last)
  d.n_keys = gkeys.empty() ? 1 : (uint32_t)gkeys.size();
  std::vector<int64_t> key_null(gkeys.size(), INT64_MIN);
  for (size_t k = 0; k < gkeys.size(); ++k) {
    d.key_type[k] = gkeys[k].c->type;
    d.key_min[k] = gkeys[k].c->min;
    d.key_max[k] = gkeys[k].c->max;
    if (gkeys[k].c->valid) {
      const int w = type_width(gkeys[k].c->type);
      const int64_t top = gkeys[k].c->type == MTL_DICT8 ? 255 : w == 4 ? INT32_MAX : INT64_MAX;
      if (gkeys[k].c->max >= top)
        return set_error(MTL_UNSUPPORTED, "GROUP BY %s: no value left for NULL",
                         gkeys[k].c->name.c_str());
      key_null[k] = d.key_max[k] = gkeys[k].c->max + 1;
    }
  }
  if (gkeys.empty()) {  // ungrouped: one constant key
    d.key_type[0] = MTL_INTEGER;
    d.key_min[0] = d.key_max[0] = 0;
  }
  if (grouped && key_order != 0) {
    // the plan's table layout depends only on the key types and ranges:
    // placeholders stand in for the columns gathered after the joins
    mtl_group_desc probe = d;
    probe.n_rows = std::max<uint64_t>(1, b.tabs[0]->rows);
    for (uint32_t k = 0; k < probe.n_keys; ++k) probe.key_col[k] = b.tabs[0];
    for (uint32_t a = 0; a < probe.n_aggs; ++a)
      if (gvals[a].c) probe.aggs[a].values = b.tabs[0];
    if (!group_is_dense(probe))
      return set_error(MTL_UNSUPPORTED, "ORDER BY group keys of a hashed group table");
  }
  res.shape = "join";
  uint64_t rows_max = 0, rows_sum = 0;
  for (auto* t : b.tabs) {
    rows_max = std::max(rows_max, t->rows);
    rows_sum += t->rows;
  }
  uint64_t ncols = proj.size() + gkeys.size() + gvals.size() + 1;
  for (auto& ax : gexpr)
    if (ax) ncols += ax->cols.size();
  res.scratch = rows_sum * 8 + rows_max * (4 * ns * 2 + 16 + 8 * ncols) +
                (grouped ? group_scratch_bytes(d) : 0) + (1u << 20);
  if (dry) return MTL_OK;
  MTL_TRY(resolve_stream(s, &s));  // NULL: the library's default stream

  // filters
  std::vector<std::unique_ptr<DevBuf>> sel(ns);
  std::vector<uint64_t> nsel(ns);
  DevBuf cnt;
  MTL_TRY(cnt.alloc(16));
  for (size_t i = 0; i < ns; ++i) {
    const CatTable& T = *b.tabs[i];
    sel[i].reset(new DevBuf());
    MTL_TRY(sel[i]->alloc(std::max<uint64_t>(T.rows, 1) * 4));
    MTL_TRY(mtl_filter_or(preds[i].data(), (uint32_t)preds[i].size(), pred_or[i], T.rows,
                          (uint32_t*)sel[i]->p, (uint32_t*)cnt.p, s));
    uint32_t n = 0;
    MTL_TRY(mtl_memcpy_d2h(&n, cnt.p, 4, s));
    if (n > T.rows) return set_error(MTL_ERR_PROCESSING, "device scan failed");
    nsel[i] = n;
  }
  // joins
  JoinRel R;
  R.ids.resize(ns);
  size_t start = 0;
  for (size_t i = 1; i < ns; ++i)
    if (nsel[i] > nsel[start]) start = i;
  R.ids[start] = std::move(sel[start]);
  R.m = nsel[start];
  for (size_t step = 1; step < ns; ++step) {
    Edge* e = nullptr;
    for (auto& x : edges)
      if (!x.used && ((R.ids[x.a] != nullptr) != (R.ids[x.c] != nullptr))) e = &x;
    if (!e) return set_error(MTL_UNSUPPORTED, "joins do not connect every table");
    e->used = true;
    int a = e->a, c = e->c;
    const CatCol *xa = e->ca, *yc = e->cc;
    std::vector<const CatCol*> ka = e->xa, kc = e->xc;
    if (!R.ids[a]) {
      std::swap(a, c);
      std::swap(xa, yc);
      std::swap(ka, kc);
    }
    // the relation's key column, gathered through its row ids; a
    // multi-column key is packed on both sides (the new table's over all its
    // rows, read through its selection)
    DevBuf kx, ky, pp, pb;
    const void* ycol = yc->dev;
    int32_t xtype = xa->type, ytype = yc->type;
    int64_t xmin = xa->min, xmax = xa->max, ymin = yc->min, ymax = yc->max;
    if (ka.size() > 1) {
      PackKeys pa{}, pc{};
      pa.n_keys = pc.n_keys = (uint32_t)ka.size();
      unsigned __int128 total = 1;
      for (int k = (int)ka.size() - 1; k >= 0; --k) {
        const int64_t lo = std::min(ka[k]->min, kc[k]->min);
        const int64_t hi = std::max(ka[k]->max, kc[k]->max);
        const unsigned __int128 span = (unsigned __int128)((uint64_t)hi - (uint64_t)lo) + 1;
        pa.col[k] = ka[k]->dev;
        pc.col[k] = kc[k]->dev;
        pa.type[k] = ka[k]->type;
        pc.type[k] = kc[k]->type;
        pa.min[k] = pc.min[k] = lo;
        pa.mult[k] = pc.mult[k] = (uint64_t)total;
        total *= span;
        if (total > ((unsigned __int128)1 << 62))
          return set_error(MTL_UNSUPPORTED, "multi-column join key ranges too wide to pack");
      }
      MTL_TRY(pack_keys(pa, (const uint32_t*)R.ids[a]->p, R.m, kx, s));
      MTL_TRY(pack_keys(pc, nullptr, b.tabs[c]->rows, ky, s));
      ycol = ky.p;
      xtype = ytype = MTL_BIGINT;
      xmin = ymin = 0;
      xmax = ymax = (int64_t)(total - 1);
    } else {
      const uint32_t wx = (uint32_t)type_width(xa->type);
      MTL_TRY(kx.alloc(std::max<uint64_t>(R.m, 1) * wx));
      if (R.m)
        MTL_TRY(mtl_gather(xa->dev, wx, (const uint32_t*)R.ids[a]->p, nullptr, R.m, kx.p, s));
    }
    uint64_t np = 0;
    // build over the new table's rows (its key unique), probe with the relation
    int32_t st = join_pairs(ycol, ytype, (const uint32_t*)sel[c]->p, nsel[c], ymin, ymax, kx.p,
                            xtype, nullptr, R.m, pp, pb, np, s, !grouped);
    if (st == MTL_UNSUPPORTED) {
      // build over the relation (its key unique), probe with the new table
      DevBuf pp2, pb2;
      st = join_pairs(kx.p, xtype, nullptr, R.m, xmin, xmax, ycol, ytype,
                      (const uint32_t*)sel[c]->p, nsel[c], pp2, pb2, np, s, !grouped);
      if (st == MTL_UNSUPPORTED) {
        // keys repeat on both sides: many-to-many, the new table sorted into runs
        st = join_pairs_multi(ycol, ytype, (const uint32_t*)sel[c]->p, nsel[c], ymin, ymax, kx.p,
                              xtype, nullptr, R.m, pp, pb, np, s, !grouped);
      } else {
        MTL_TRY(st);
        std::swap(pp.p, pb2.p);
        std::swap(pp.n, pb2.n);
        std::swap(pb.p, pp2.p);
        std::swap(pb.n, pp2.n);
      }
    }
    MTL_TRY(st);
    // pp: relation positions, pb: rows of table c (in the order of the
    // probing side)
    for (size_t i = 0; i < ns; ++i) {
      if (!R.ids[i]) continue;
      std::unique_ptr<DevBuf> g(new DevBuf());
      MTL_TRY(gather_ids(R.ids[i]->p, pp.p, np, *g, s));
      R.ids[i] = std::move(g);
    }
    R.ids[c].reset(new DevBuf());
    std::swap(R.ids[c]->p, pb.p);
    std::swap(R.ids[c]->n, pb.n);
    R.m = np;
  }
  const uint64_t m = R.m;
  // the validity bytes of a column with NULLs, gathered into relation order
  auto gather_valid = [&](const Col& c, DevBuf& out) -> int32_t {
    MTL_TRY(out.alloc(std::max<uint64_t>(m, 1)));
    if (m) MTL_TRY(mtl_gather(c.c->valid, 1, (const uint32_t*)R.ids[c.src]->p, nullptr, m, out.p, s));
    return MTL_OK;
  };
  if (!grouped) {
    DevBuf perm, out;
    uint64_t n_out = m;
    const void* order = nullptr;  // positions in output order (NULL: relation order)
    if (okey.c) {
      DevBuf kcol;
      const uint32_t w = (uint32_t)type_width(okey.c->type);
      MTL_TRY(kcol.alloc(std::max<uint64_t>(m, 1) * w));
      if (m) MTL_TRY(mtl_gather(okey.c->dev, w, (const uint32_t*)R.ids[okey.src]->p, nullptr, m,
                                kcol.p, s));
      MTL_TRY(perm.alloc(std::max<uint64_t>(m, 1) * 4));
      uint32_t n = 0;
      if (okey.c->valid && m) {
        // NULLs last: the valid positions sorted, then the NULL positions
        DevBuf vb, vpos, npos, vcnt;
        MTL_TRY(gather_valid(okey, vb));
        MTL_TRY(vpos.alloc(m * 4));
        MTL_TRY(npos.alloc(m * 4));
        MTL_TRY(vcnt.alloc(16));
        mtl_pred pv{vb.p, MTL_DICT8, MTL_CMP_EQ, 1}, pn{vb.p, MTL_DICT8, MTL_CMP_EQ, 0};
        MTL_TRY(mtl_filter(&pv, 1, m, nullptr, nullptr, m, (uint32_t*)vpos.p, (uint32_t*)vcnt.p, s));
        MTL_TRY(mtl_filter(&pn, 1, m, nullptr, nullptr, m, (uint32_t*)npos.p,
                           (uint32_t*)vcnt.p + 1, s));
        uint32_t c2[2] = {0, 0};
        MTL_TRY(mtl_memcpy_d2h(c2, vcnt.p, 8, s));
        if (c2[0] > m || c2[1] > m - c2[0]) return set_error(MTL_ERR_PROCESSING, "device scan failed");
        MTL_TRY(mtl_sort(kcol.p, okey.c->type, (const uint32_t*)vpos.p, (const uint32_t*)vcnt.p, m,
                         odesc ? 1 : 0, (uint32_t*)perm.p, &n, s));
        if (c2[1])
          MTL_HIP(hipMemcpyAsync((uint32_t*)perm.p + n, npos.p, (uint64_t)c2[1] * 4,
                                 hipMemcpyDeviceToDevice, s->stream));
      } else {
        MTL_TRY(mtl_sort(kcol.p, okey.c->type, nullptr, nullptr, m, odesc ? 1 : 0,
                         (uint32_t*)perm.p, &n, s));
      }
      order = perm.p;
    }
    if (q.limit >= 0 && (uint64_t)q.limit < n_out) n_out = (uint64_t)q.limit;
    res.rows = n_out;
    MTL_TRY(out.alloc(std::max<uint64_t>(n_out, 1) * 8));
    for (size_t i = 0; i < proj.size(); ++i) {
      mtl_result::Column col;
      col.name = proj_name[i];
      col.type = proj[i].c->type;
      col.scale = proj[i].c->scale;
      col.dict = proj[i].c->dict;
      const uint32_t w = (uint32_t)type_width(col.type);
      const void* rows = R.ids[proj[i].src]->p;
      DevBuf idx;  // the column's row ids in output order
      if (order && n_out) {
        MTL_TRY(gather_ids(rows, order, n_out, idx, s));
        rows = idx.p;
      }
      if (n_out) MTL_TRY(mtl_gather(proj[i].c->dev, w, (const uint32_t*)rows, nullptr, n_out, out.p, s));
      MTL_TRY(download(col.data, out.p, n_out * w, s));
      if (proj[i].c->valid) {  // a projection of NULL is NULL
        if (n_out) MTL_TRY(mtl_gather(proj[i].c->valid, 1, (const uint32_t*)rows, nullptr, n_out,
                                      out.p, s));
        MTL_TRY(download(col.valid, out.p, n_out, s));
      }
      res.cols.push_back(std::move(col));
    }
    return MTL_OK;
  }
  // aggregates over the gathered key and value columns
  if (gkeys.empty() && m == 0) {
    // no rows: one row, COUNTs 0 and every other aggregate NULL
    res.rows = 1;
    for (auto& o : gouts) {
      mtl_result::Column col;
      col.name = o.name;
      const int a = o.agg;
      if (agg_func[a] == MTL_AGG_COUNT || o.mode == 1) {
        col.type = MTL_BIGINT;
        col.data.assign(8, 0);
      } else {
        agg_out_type(o.mode == 2 ? MTL_AGG_AVG : agg_func[a], agg_type[a], agg_scale[a], col.type,
                     col.scale);
        col.data.assign((size_t)type_width(col.type), 0);
        col.valid.assign(1, 0);
      }
      res.cols.push_back(std::move(col));
    }
    return MTL_OK;
  }
  std::vector<std::unique_ptr<DevBuf>> mats;
  auto gathered = [&](const Col& c, const void** out) -> int32_t {
    const uint32_t w = (uint32_t)type_width(c.c->type);
    mats.emplace_back(new DevBuf());
    MTL_TRY(mats.back()->alloc(std::max<uint64_t>(m, 1) * w));
    if (m) MTL_TRY(mtl_gather(c.c->dev, w, (const uint32_t*)R.ids[c.src]->p, nullptr, m,
                              mats.back()->p, s));
    *out = mats.back()->p;
    return MTL_OK;
  };
  d.n_rows = m;
  if (gkeys.empty()) {
    mats.emplace_back(new DevBuf());
    MTL_TRY(mats.back()->alloc(std::max<uint64_t>(m, 1) * 4));
    MTL_HIP(hipMemsetAsync(mats.back()->p, 0, std::max<uint64_t>(m, 1) * 4, s->stream));
    d.key_col[0] = mats.back()->p;
  }
  for (size_t k = 0; k < gkeys.size(); ++k) {
    MTL_TRY(gathered(gkeys[k], &d.key_col[k]));
    if (key_null[k] != INT64_MIN && m) {  // NULL keys -> the value above the range
      DevBuf vb;
      MTL_TRY(gather_valid(gkeys[k], vb));
      hipLaunchKernelGGL(k_null_fill, dim3(grid_for(m)), dim3(256), 0, s->stream,
                         (void*)d.key_col[k], (uint32_t)type_width(gkeys[k].c->type),
                         (const uint8_t*)vb.p, m, (uint64_t)key_null[k]);
      MTL_HIP(hipGetLastError());
    }
  }
  // validity of each NULL-dependent aggregate: the AND over its columns
  auto agg_valid = [&](size_t a, DevBuf& vb) -> int32_t {
    MTL_TRY(gather_valid(gnull[a][0], vb));
    for (size_t k = 1; k < gnull[a].size(); ++k) {
      DevBuf v2;
      MTL_TRY(gather_valid(gnull[a][k], v2));
      if (m) {
        hipLaunchKernelGGL(k_valid_and, dim3(grid_for(m)), dim3(256), 0, s->stream,
                           (uint8_t*)vb.p, (const uint8_t*)v2.p, m);
        MTL_HIP(hipGetLastError());
      }
    }
    return MTL_OK;
  };
  DevBuf eflag;
  bool any_expr = false;
  for (size_t a = 0; a < gvals.size(); ++a) {
    if (!gvals[a].c) continue;
    if (gcount[a]) {  // COUNT of the valid values: SUM of a 0/1 int32 column
      DevBuf vb;
      MTL_TRY(agg_valid(a, vb));
      mats.emplace_back(new DevBuf());
      MTL_TRY(mats.back()->alloc(std::max<uint64_t>(m, 1) * 4));
      if (m) {
        hipLaunchKernelGGL(k_valid_i32, dim3(grid_for(m)), dim3(256), 0, s->stream,
                           (const uint8_t*)vb.p, m, (int32_t*)mats.back()->p);
        MTL_HIP(hipGetLastError());
      }
      d.aggs[a].values = mats.back()->p;
      continue;
    }
    if (!gexpr[a]) {
      MTL_TRY(gathered(gvals[a], &d.aggs[a].values));
      continue;
    }
    // gather the expression's columns into row order, then evaluate it
    AggExpr& ax = *gexpr[a];
    for (size_t k = 0; k < ax.cols.size(); ++k) MTL_TRY(gathered(ax.cols[k], &ax.x.col[k]));
    if (!any_expr) {
      MTL_TRY(eflag.alloc(4));
      MTL_HIP(hipMemsetAsync(eflag.p, 0, 4, s->stream));
      any_expr = true;
    }
    mats.emplace_back(new DevBuf());
    MTL_TRY(mats.back()->alloc(std::max<uint64_t>(m, 1) * 8));
    MTL_TRY(mtl::expr_eval_i64(ax.x, m, (int64_t*)mats.back()->p, (uint32_t*)eflag.p, s->stream));
    d.aggs[a].values = mats.back()->p;
  }
  // NULL values of the other aggregates -> the aggregate's identity
  for (size_t a = 0; a < gvals.size(); ++a) {
    if (!gvals[a].c || gcount[a] || gnull[a].empty() || !m) continue;
    const int32_t t = agg_type[a];
    const uint32_t w = gexpr[a] ? 8u : (uint32_t)type_width(t);
    uint64_t fill = 0;  // SUM (and AVG, a SUM here)
    const bool fl = t == MTL_FLOAT || t == MTL_DOUBLE;
    if (agg_func[a] == MTL_AGG_MIN)  // NaN sorts above every value, as in DuckDB
      fill = fl ? (w == 8 ? 0x7FF8000000000000ull : 0x7FC00000ull)
                : (w == 8 ? (uint64_t)INT64_MAX : (uint64_t)INT32_MAX);
    else if (agg_func[a] == MTL_AGG_MAX)
      fill = fl ? (w == 8 ? 0xFFF0000000000000ull : 0xFF800000ull)
                : (w == 8 ? (uint64_t)INT64_MIN : (uint64_t)(uint32_t)INT32_MIN);
    DevBuf vb;
    MTL_TRY(agg_valid(a, vb));
    hipLaunchKernelGGL(k_null_fill, dim3(grid_for(m)), dim3(256), 0, s->stream,
                       (void*)d.aggs[a].values, w, (const uint8_t*)vb.p, m, fill);
    MTL_HIP(hipGetLastError());
  }
  if (any_expr) {
    uint32_t ef = 0;
    MTL_TRY(mtl_memcpy_d2h(&ef, eflag.p, 4, s));
    if (ef) return set_error(MTL_ERR_OVERFLOW, "Overflow in aggregate expression");
  }
  std::vector<const CatCol*> ek = gkey_cols;
  GroupEmit ge{ek, agg_func, agg_type, agg_scale, order_agg, odesc, key_order, then_keys,
               key_null};
  return emit_groups(q, d, ge, gouts, res, s);
}

// ---------------------------------------------------------------------------
// Driver
// ---------------------------------------------------------------------------
static int32_t plan_and_run(const char* sql, mtl_result& res, mtl_stream* s, bool dry) {
  if (!sql) return set_error(MTL_ERR_INVALID_ARGUMENT, "sql is NULL");
  std::vector<Tok> toks;
  std::string err;
  if (!lex(sql, toks, err)) return set_error(MTL_UNSUPPORTED, "%s", err.c_str());
  Parser ps(toks);
  Query q;
  if (!ps.query(q)) return set_error(MTL_UNSUPPORTED, "%s", ps.err.c_str());
  std::lock_guard<std::mutex> lk(g_cat_mu);
  Bound b;
  for (auto& f : q.from) {
    auto it = catalog().find(f.first);
    if (it == catalog().end())
      return set_error(MTL_UNSUPPORTED, "table %s is not registered", f.first.c_str());
    b.tabs.push_back(&it->second);
    b.names.push_back(f.second.empty() ? f.first : f.second);
  }
  bool has_agg = false;
  for (auto& it : q.items) has_agg |= it.e->k == Expr::AGG;
  if (b.tabs.size() == 1 && q.joins.empty()) {
    if (q.group.empty() && !has_agg) return run_scan(q, b, res, s, dry);
    if (!q.group.empty()) {
      // ORDER BY the group keys is the generic path's
      const int32_t st = run_group(q, b, res, s, dry);
      if (st != MTL_UNSUPPORTED) return st;
      res = mtl_result{};
    }
    return run_join(q, b, res, s, dry);  // aggregates without GROUP BY
  }
  if ((b.tabs.size() == 2 || b.tabs.size() == 3) && !q.group.empty() && has_agg) {
    // the star pipeline when the query has its shape (grouped by the parent
    // key, one SUM, top-K), else the generic joins
    const int32_t st = run_star(q, b, res, s, dry);
    if (st != MTL_UNSUPPORTED) return st;
    res = mtl_result{};
  }
  return run_join(q, b, res, s, dry);
}

// ---------------------------------------------------------------------------
// gpu_rewrite_query: print the query back with its predicates and scalar
// projections written as the gpu_* hook calls.
// ---------------------------------------------------------------------------
static std::string lit_sql(const Lit& l) {
  switch (l.k) {
    case Lit::INT:
    case Lit::FLT: return l.text;
    case Lit::DATE: return "DATE '" + l.s + "'";
    case Lit::TS: return "TIMESTAMP '" + l.s + "'";
    default: {
      std::string o = "'";
      for (char c : l.s) o += c == '\'' ? std::string("''") : std::string(1, c);
      return o + "'";
    }
  }
}
static std::string col_sql(const ColRef& c) { return c.rel.empty() ? c.name : c.rel + "." + c.name; }

static std::string expr_sql(const ExprP& e) {
  ColRef c;
  int op;
  Lit l;
  if (as_projection(e, c, op, l) && (op == MTL_PROJ_ADD || op == MTL_PROJ_MUL))
    return std::string(op == MTL_PROJ_ADD ? "gpu_add_scalar(" : "gpu_mul_scalar(") + col_sql(c) +
           ", " + lit_sql(l) + ")";
  switch (e->k) {
    case Expr::COL: return col_sql(e->col);
    case Expr::LIT: return lit_sql(e->lit);
    case Expr::STAR: return "*";
    case Expr::AGG: return e->fn + "(" + (e->a ? expr_sql(e->a) : std::string("*")) + ")";
    case Expr::CALL: return e->fn + "(" + expr_sql(e->a) + ", " + expr_sql(e->b) + ")";
    default: {
      const char* o = e->k == Expr::ADD ? " + " : e->k == Expr::SUB ? " - " : " * ";
      return "(" + expr_sql(e->a) + o + expr_sql(e->b) + ")";
    }
  }
}

static int32_t rewrite(const char* sql, std::string& out) {
  if (!sql) return set_error(MTL_ERR_INVALID_ARGUMENT, "sql is NULL");
  std::vector<Tok> toks;
  std::string err;
  if (!lex(sql, toks, err)) return set_error(MTL_UNSUPPORTED, "%s", err.c_str());
  Parser ps(toks);
  Query q;
  if (!ps.query(q)) return set_error(MTL_UNSUPPORTED, "%s", ps.err.c_str());
  static const char* hook[] = {"gpu_eq_scalar", "gpu_lt_scalar", "gpu_le_scalar",
                               "gpu_gt_scalar", "gpu_ge_scalar"};
  out = "SELECT ";
  for (size_t i = 0; i < q.items.size(); ++i) {
    if (i) out += ", ";
    out += expr_sql(q.items[i].e);
    if (!q.items[i].alias.empty()) out += " AS " + q.items[i].alias;
  }
  out += " FROM ";
  for (size_t i = 0; i < q.from.size(); ++i) {
    if (i) out += ", ";
    out += q.from[i].first;
    if (!q.from[i].second.empty()) out += " " + q.from[i].second;
  }
  std::vector<std::string> conj;
  for (auto& j : q.joins) conj.push_back(col_sql(j.l) + " = " + col_sql(j.r));
  std::string run;  // the current run of OR-ed predicates
  size_t in_run = 0;
  for (auto& p : q.preds) {
    if (p.nulls)
      run += (in_run ? " OR " : "") + col_sql(p.col) + (p.nulls == 1 ? " IS NULL" : " IS NOT NULL");
    else
      run += (in_run ? " OR " : "") + std::string(hook[p.op]) + "(" + col_sql(p.col) + ", " +
             lit_sql(p.lit) + ")";
    ++in_run;
    if (!p.or_next) {
      conj.push_back(in_run > 1 ? "(" + run + ")" : run);
      run.clear();
      in_run = 0;
    }
  }
  for (size_t i = 0; i < conj.size(); ++i) out += (i ? " AND " : " WHERE ") + conj[i];
  for (size_t i = 0; i < q.group.size(); ++i) out += (i ? ", " : " GROUP BY ") + col_sql(q.group[i]);
  for (size_t i = 0; i < q.order.size(); ++i)
    out += (i ? ", " : " ORDER BY ") + expr_sql(q.order[i].e) + (q.order[i].desc ? " DESC" : "");
  if (q.limit >= 0) out += " LIMIT " + std::to_string(q.limit);
  return MTL_OK;
}

}  // namespace sql
}  // namespace mtl

namespace mtl {
void catalog_drop_table(const char* table) {
  using namespace mtl::sql;
  std::lock_guard<std::mutex> lk(g_cat_mu);
  if (table) catalog().erase(lower(table));
  else catalog().clear();
}

// Admission size of a statement for the serving loop: its plan's device
// scratch (MTL_UNSUPPORTED and the reason when it does not lower).
int32_t sql_scratch_bytes(const char* sql, uint64_t* bytes) {
  if (!sql) return set_error(MTL_ERR_INVALID_ARGUMENT, "sql is NULL");
  mtl_result res;
  MTL_TRY(sql::plan_and_run(sql, res, nullptr, true));
  *bytes = res.scratch;
  return MTL_OK;
}
}  // namespace mtl

extern "C" {

int32_t mtl_catalog_add_table(const char* table, uint64_t n_rows) {
  using namespace mtl::sql;
  if (!table) return mtl::set_error(MTL_ERR_INVALID_ARGUMENT, "table is NULL");
  std::lock_guard<std::mutex> lk(g_cat_mu);
  CatTable t;
  t.name = lower(table);
  t.rows = n_rows;
  catalog()[t.name] = t;
  return MTL_OK;
}

int32_t mtl_catalog_add_column(const char* table, const char* column, int32_t type,
                               int32_t scale, const void* dev, int64_t min_value,
                               int64_t max_value, int32_t unique) {
  using namespace mtl::sql;
  if (!table || !column || !dev) return mtl::set_error(MTL_ERR_INVALID_ARGUMENT, "NULL argument");
  if (!mtl::type_width(type) || type == MTL_HUGEINT)
    return mtl::set_error(MTL_ERR_INVALID_ARGUMENT, "column type %d", type);
  std::lock_guard<std::mutex> lk(g_cat_mu);
  auto it = catalog().find(lower(table));
  if (it == catalog().end())
    return mtl::set_error(MTL_ERR_INVALID_ARGUMENT, "table %s is not registered", table);
  CatCol c;
  c.name = lower(column);
  c.type = type;
  c.scale = scale;
  c.dev = dev;
  c.min = min_value;
  c.max = max_value;
  c.unique = unique != 0;
  for (auto& old : it->second.cols)
    if (old.name == c.name) {
      old = c;
      return MTL_OK;
    }
  it->second.cols.push_back(c);
  return MTL_OK;
}

int32_t mtl_catalog_set_dictionary(const char* table, const char* column,
                                   const char* const* values, uint32_t n) {
  using namespace mtl::sql;
  if (!table || !column || (n && !values))
    return mtl::set_error(MTL_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(g_cat_mu);
  auto it = catalog().find(lower(table));
  if (it == catalog().end())
    return mtl::set_error(MTL_ERR_INVALID_ARGUMENT, "table %s is not registered", table);
  for (auto& c : it->second.cols)
    if (c.name == lower(column)) {
      c.dict.assign(values, values + n);
      return MTL_OK;
    }
  return mtl::set_error(MTL_ERR_INVALID_ARGUMENT, "column %s is not registered", column);
}

int32_t mtl_catalog_set_validity(const char* table, const char* column, const uint8_t* valid) {
  using namespace mtl::sql;
  if (!table || !column) return mtl::set_error(MTL_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(g_cat_mu);
  auto it = catalog().find(lower(table));
  if (it == catalog().end())
    return mtl::set_error(MTL_ERR_INVALID_ARGUMENT, "table %s is not registered", table);
  for (auto& c : it->second.cols)
    if (c.name == lower(column)) {
      c.valid = valid;
      return MTL_OK;
    }
  return mtl::set_error(MTL_ERR_INVALID_ARGUMENT, "column %s is not registered", column);
}

int32_t mtl_catalog_drop(const char* table) {
  mtl::catalog_drop_table(table);
  mtl::ingest_release(table);  // after the catalog lock: ingestion takes the locks the other way
  return MTL_OK;
}


int32_t mtl_sql_explain(const char* sql, char* buf, size_t len) {
  mtl_result res;
  const int32_t st = mtl::sql::plan_and_run(sql, res, nullptr, true);
  if (buf && len) std::snprintf(buf, len, "%s", st == MTL_OK ? res.shape.c_str() : mtl::last_error());
  return st;
}

int32_t mtl_sql_execute(const char* sql, mtl_result** out, mtl_stream* s) {
  if (!out) return mtl::set_error(MTL_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  MTL_TRY(mtl::ensure_device());
  auto* res = new mtl_result();
  const int32_t st = mtl::sql::plan_and_run(sql, *res, s, false);
  if (st != MTL_OK) {
    delete res;
    return st;
  }
  *out = res;
  return MTL_OK;
}

int32_t mtl_result_shape(const mtl_result* r, uint32_t* n_cols, uint64_t* n_rows) {
  if (!r) return mtl::set_error(MTL_ERR_INVALID_ARGUMENT, "result is NULL");
  if (n_cols) *n_cols = (uint32_t)r->cols.size();
  if (n_rows) *n_rows = r->rows;
  return MTL_OK;
}

int32_t mtl_result_column(const mtl_result* r, uint32_t i, const char** name, int32_t* type,
                          int32_t* scale, const void** data) {
  if (!r || i >= r->cols.size())
    return mtl::set_error(MTL_ERR_INVALID_ARGUMENT, "no result column %u", i);
  const auto& c = r->cols[i];
  if (name) *name = c.name.c_str();
  if (type) *type = c.type;
  if (scale) *scale = c.scale;
  if (data) *data = c.data.data();
  return MTL_OK;
}

int32_t mtl_result_validity(const mtl_result* r, uint32_t i, const uint8_t** valid) {
  if (!r || i >= r->cols.size() || !valid)
    return mtl::set_error(MTL_ERR_INVALID_ARGUMENT, "no result column %u", i);
  *valid = r->cols[i].valid.empty() ? nullptr : r->cols[i].valid.data();
  return MTL_OK;
}

int32_t mtl_result_free(mtl_result* r) {
  delete r;
  return MTL_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Arrow hand-back: a result as one record batch in the Arrow C data
// interface (a struct array with one child per column).  Every array and
// schema owns its buffers through private_data and frees them in its
// release callback, so the consumer (DuckDB's Arrow scan, pyarrow) keeps
// the batch independent of the mtl_result.
// ---------------------------------------------------------------------------
namespace {

struct ArrowOwn {
  std::vector<std::vector<uint8_t>> bufs;
  std::vector<const void*> ptrs;
  std::vector<ArrowArray*> kids;
  ArrowArray* dict = nullptr;
};

struct SchemaOwn {
  std::string format, name;
  std::vector<ArrowSchema*> kids;
  ArrowSchema* dict = nullptr;
};

void release_array(ArrowArray* a) {
  if (!a || !a->release) return;
  auto* own = (ArrowOwn*)a->private_data;
  for (ArrowArray* k : own->kids) {
    if (k->release) k->release(k);
    delete k;
  }
  if (own->dict) {
    if (own->dict->release) own->dict->release(own->dict);
    delete own->dict;
  }
  delete own;
  a->release = nullptr;
}

void release_schema(ArrowSchema* s) {
  if (!s || !s->release) return;
  auto* own = (SchemaOwn*)s->private_data;
  for (ArrowSchema* k : own->kids) {
    if (k->release) k->release(k);
    delete k;
  }
  if (own->dict) {
    if (own->dict->release) own->dict->release(own->dict);
    delete own->dict;
  }
  delete own;
  s->release = nullptr;
}

void make_schema(ArrowSchema* s, const std::string& format, const std::string& name,
                 std::vector<ArrowSchema*> kids, ArrowSchema* dict, bool nullable = false) {
  auto* own = new SchemaOwn{format, name, std::move(kids), dict};
  s->format = own->format.c_str();
  s->name = own->name.c_str();
  s->metadata = nullptr;
  s->flags = nullable ? 2 : 0;  // ARROW_FLAG_NULLABLE
  s->n_children = (int64_t)own->kids.size();
  s->children = own->kids.empty() ? nullptr : own->kids.data();
  s->dictionary = dict;
  s->release = release_schema;
  s->private_data = own;
}

// buffers: the validity bitmap (absent when `valid` is empty: no NULLs)
// then `bufs`; valid holds one byte per row (1 = valid).
void make_array(ArrowArray* a, int64_t length, std::vector<std::vector<uint8_t>> bufs,
                std::vector<ArrowArray*> kids, ArrowArray* dict,
                const std::vector<uint8_t>& valid = {}) {
  auto* own = new ArrowOwn();
  int64_t nulls = 0;
  std::vector<uint8_t> bitmap;
  if (!valid.empty()) {
    bitmap.assign((size_t)(length + 7) / 8, 0);
    for (int64_t i = 0; i < length; ++i) {
      if (valid[i]) bitmap[i / 8] |= (uint8_t)(1u << (i % 8));  // LSB first
      else ++nulls;
    }
  }
  own->bufs = std::move(bufs);
  if (!valid.empty()) own->bufs.push_back(std::move(bitmap));
  own->ptrs.push_back(valid.empty() ? nullptr : own->bufs.back().data());
  for (size_t i = 0; i + (valid.empty() ? 0 : 1) < own->bufs.size(); ++i)
    own->ptrs.push_back(own->bufs[i].data());
  own->kids = std::move(kids);
  own->dict = dict;
  a->length = length;
  a->null_count = nulls;
  a->offset = 0;
  a->n_buffers = (int64_t)own->ptrs.size();
  a->n_children = (int64_t)own->kids.size();
  a->buffers = own->ptrs.data();
  a->children = own->kids.empty() ? nullptr : own->kids.data();
  a->dictionary = dict;
  a->release = release_array;
  a->private_data = own;
}

// A utf8 array of the dictionary's values.
void make_utf8(ArrowSchema* s, ArrowArray* a, const std::vector<std::string>& vals) {
  std::vector<uint8_t> offs((vals.size() + 1) * 4), chars;
  int32_t at = 0;
  std::memcpy(offs.data(), &at, 4);
  for (size_t i = 0; i < vals.size(); ++i) {
    chars.insert(chars.end(), vals[i].begin(), vals[i].end());
    at += (int32_t)vals[i].size();
    std::memcpy(offs.data() + (i + 1) * 4, &at, 4);
  }
  make_schema(s, "u", "", {}, nullptr);
  make_array(a, (int64_t)vals.size(), {std::move(offs), std::move(chars)}, {}, nullptr);
}

}  // namespace

extern "C" {

int32_t mtl_result_export_arrow(const mtl_result* r, ArrowSchema* schema, ArrowArray* array) {
  using namespace mtl;
  if (!r || !schema || !array) return set_error(MTL_ERR_INVALID_ARGUMENT, "NULL argument");
  const int64_t n = (int64_t)r->rows;
  std::vector<ArrowSchema*> cs;
  std::vector<ArrowArray*> ca;
  auto fail = [&](const char* why, int32_t t) {
    for (auto* s : cs) {
      s->release(s);
      delete s;
    }
    for (auto* a : ca) {
      a->release(a);
      delete a;
    }
    return set_error(MTL_UNSUPPORTED, "%s (type %d)", why, t);
  };
  for (const auto& c : r->cols) {
    auto* s = new ArrowSchema();
    auto* a = new ArrowArray();
    const int w = type_width(c.type);
    if ((uint64_t)c.data.size() < (uint64_t)n * (uint64_t)w ||
        (!c.valid.empty() && c.valid.size() < (uint64_t)n)) {
      delete s;
      delete a;
      return fail("result column shorter than its rows", c.type);
    }
    std::vector<uint8_t> data(c.data.begin(), c.data.begin() + (size_t)n * w);
    std::string fmt;
    switch (c.type) {
      case MTL_INTEGER: fmt = "i"; break;
      case MTL_BIGINT: fmt = "l"; break;
      case MTL_FLOAT: fmt = "f"; break;
      case MTL_DOUBLE: fmt = "g"; break;
      case MTL_DATE: fmt = "tdD"; break;
      case MTL_TIMESTAMP: fmt = "tsu:"; break;
      case MTL_DECIMAL:
      case MTL_HUGEINT: {
        // decimal128 (DuckDB's Arrow type for DECIMAL and HUGEINT): int64
        // values sign-extended to 16 bytes, HUGEINT {lower, upper} as is
        fmt = std::string("d:") + (c.type == MTL_DECIMAL ? "18," : "38,") +
              std::to_string(c.scale);
        if (c.type == MTL_DECIMAL) {
          std::vector<uint8_t> wide((size_t)n * 16);
          for (int64_t i = 0; i < n; ++i) {
            int64_t v;
            std::memcpy(&v, data.data() + i * 8, 8);
            const int64_t hi = v < 0 ? -1 : 0;
            std::memcpy(wide.data() + i * 16, &v, 8);
            std::memcpy(wide.data() + i * 16 + 8, &hi, 8);
          }
          data.swap(wide);
        }
        break;
      }
      case MTL_DICT8:
      case MTL_DICT32: {
        // dictionary<uint8 | int32, utf8>; codes without a dictionary as
        // plain integers
        if (c.dict.empty()) {
          fmt = c.type == MTL_DICT8 ? "C" : "i";
          break;
        }
        for (int64_t i = 0; i < n; ++i) {
          uint32_t code = 0;
          if (c.type == MTL_DICT8) code = data[i];
          else std::memcpy(&code, data.data() + i * 4, 4);
          if (code >= c.dict.size()) {
            delete s;
            delete a;
            return fail("dictionary code outside the dictionary", c.type);
          }
        }
        auto* ds = new ArrowSchema();
        auto* da = new ArrowArray();
        make_utf8(ds, da, c.dict);
        make_schema(s, c.type == MTL_DICT8 ? "C" : "i", c.name, {}, ds, !c.valid.empty());
        make_array(a, n, {std::move(data)}, {}, da, c.valid);
        cs.push_back(s);
        ca.push_back(a);
        continue;
      }
      default:
        delete s;
        delete a;
        return fail("no Arrow export for the result type", c.type);
    }
    make_schema(s, fmt, c.name, {}, nullptr, !c.valid.empty());
    make_array(a, n, {std::move(data)}, {}, nullptr, c.valid);
    cs.push_back(s);
    ca.push_back(a);
  }
  make_schema(schema, "+s", "", std::move(cs), nullptr);
  make_array(array, n, {}, std::move(ca), nullptr);
  return MTL_OK;
}

namespace {
// Per-thread buffers of the scalar hooks: a stream, device memory and
// pinned staging, grown on demand and reused by every call of the thread
// (DuckDB calls a hook once per vector of 2048 values), so a call makes no
// allocation and synchronises once.  They are held until the process ends.
struct HookBufs {
  mtl_stream* s = nullptr;
  void* dev = nullptr;
  uint64_t dev_cap = 0;
  void* host = nullptr;
  uint64_t host_cap = 0;
};
thread_local HookBufs t_hook;

int32_t hook_reserve(uint64_t dev_bytes, uint64_t host_bytes) {
  using namespace mtl;
  HookBufs& h = t_hook;
  if (!h.s) MTL_TRY(mtl_stream_create(&h.s));
  if (dev_bytes > h.dev_cap) {
    uint64_t cap = 1 << 16;
    while (cap < dev_bytes) cap *= 2;
    void* p = nullptr;
    MTL_TRY(dev_alloc(cap, &p, "scalar hook buffer"));
    dev_free(h.dev, h.dev_cap);
    h.dev = p;
    h.dev_cap = cap;
  }
  if (host_bytes > h.host_cap) {
    uint64_t cap = 1 << 16;
    while (cap < host_bytes) cap *= 2;
    void* p = nullptr;
    MTL_HIP(hipHostMalloc(&p, cap, hipHostMallocDefault));
    if (h.host) hipHostFree(h.host);
    h.host = p;
    h.host_cap = cap;
  }
  return MTL_OK;
}
}  // namespace

int32_t mtl_scalar_hook(int32_t hook, int32_t type, const void* values, uint64_t n,
                        int64_t scalar_bits, void* out) {
  using namespace mtl;
  if (hook < MTL_HOOK_ADD || hook > MTL_HOOK_GE)
    return set_error(MTL_ERR_INVALID_ARGUMENT, "hook %d", hook);
  if ((!values || !out) && n) return set_error(MTL_ERR_INVALID_ARGUMENT, "NULL argument");
  const int w = type_width(type);
  if (!w || type == MTL_HUGEINT) return set_error(MTL_UNSUPPORTED, "hook on type %d", type);
  if (n == 0) return MTL_OK;
  MTL_TRY(ensure_device());
  const bool proj = hook == MTL_HOOK_ADD || hook == MTL_HOOK_MUL;
  const uint64_t in_b = (n * w + 255) & ~uint64_t(255);
  const uint64_t out_w = proj ? (uint64_t)w : 1u;
  const uint64_t res_b = (n * out_w + 255) & ~uint64_t(255);
  MTL_TRY(hook_reserve(in_b + res_b + 256, in_b + res_b + 256));
  HookBufs& h = t_hook;
  hipStream_t st = h.s->stream;
  char* din = (char*)h.dev;
  char* dres = din + in_b;
  uint32_t* dflag = (uint32_t*)(dres + res_b);
  char* hin = (char*)h.host;
  char* hres = hin + in_b;
  const uint32_t* hflag = (const uint32_t*)(hres + res_b);
  std::memcpy(hin, values, n * w);
  MTL_HIP(hipMemcpyAsync(din, hin, n * w, hipMemcpyHostToDevice, st));
  if (proj) {
    ZeroList zl;
    zl.add(dflag, 4);
    MTL_TRY(launch_zero(zl, st));
    MTL_TRY(mtl_project(din, type, nullptr, nullptr, n,
                        hook == MTL_HOOK_ADD ? MTL_PROJ_ADD : MTL_PROJ_MUL, scalar_bits, dres,
                        dflag, h.s));
    MTL_HIP(hipMemcpyAsync(hres, dres, res_b + 16, hipMemcpyDeviceToHost, st));
  } else {
    mtl_pred p{};
    p.column = din;
    p.type = type;
    p.op = hook - MTL_HOOK_EQ;  // EQ LT LE GT GE in compare-op order
    p.scalar = scalar_bits;
    MTL_TRY(mtl_compare(&p, 1, n, (uint8_t*)dres, h.s));
    MTL_HIP(hipMemcpyAsync(hres, dres, n, hipMemcpyDeviceToHost, st));
  }
  MTL_HIP(stream_wait(st));
  if (proj && hflag[0])
    return set_error(MTL_ERR_OVERFLOW, "Overflow in %s",
                     hook == MTL_HOOK_ADD ? "addition" : "multiplication");
  std::memcpy(out, hres, n * out_w);
  return MTL_OK;
}

int32_t mtl_rewrite_query(const char* sql, char* buf, size_t len, size_t* needed) {
  std::string out;
  const int32_t st = mtl::sql::rewrite(sql, out);
  if (st != MTL_OK) return st;
  if (needed) *needed = out.size() + 1;
  if (!buf || len <= out.size())
    return mtl::set_error(MTL_ERR_INVALID_ARGUMENT, "buffer of %zu bytes too small", len);
  std::memcpy(buf, out.c_str(), out.size() + 1);
  return MTL_OK;
}

}  // extern "C"
