// tree.cpp -- host-side expression engine (see tree.h).
#include "tree.h"

#include <charconv>

#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>

#include "../../include/rr_fitness.h"

namespace rrh {

static const double EPS = 1e-12;   // node.h:13-14, pow(10, -12)
static inline bool value_zero(double v) { return std::fabs(v) < EPS; }   // node.h:333-335
static inline bool value_one(double v) { return std::fabs(v - 1) < EPS; } // node.h:337-339

void subtree_sizes(const Flat& f, std::vector<int32_t>& size) {
  const int m = (int)f.size();
  size.resize(m);
  std::vector<int32_t> st;
  st.reserve(64);
  for (int i = m - 1; i >= 0; --i) {
    int32_t s = 1;
    for (int a = arity(f[i].t); a > 0; --a) {
      if (st.empty()) throw std::runtime_error("malformed prefix tree");
      s += st.back();
      st.pop_back();
    }
    size[i] = s;
    st.push_back(s);
  }
  if (st.size() != 1) throw std::runtime_error("malformed prefix tree");
}

// std::to_string(int) / std::to_string(double) as used by node::to_string;
// (int)round(c) outside the int range is what x86 cvttsd2si returns (INT_MIN).
void append_const(double c, std::string& out) {
  // std::to_chars in fixed notation with precision 6 writes what
  // printf("%f") writes (the standard defines it so), without printf's cost
  char buf[400];   // "-" + 309 integer digits + "." + 6 decimals fits
  const double rc = std::round(c);
  if (std::fabs(rc - c) < EPS) {
    const int iv = (rc >= -2147483648.0 && rc < 2147483648.0) ? (int)rc : INT_MIN;
    const std::to_chars_result r = std::to_chars(buf, buf + sizeof buf, iv);
    out.append(buf, r.ptr);
    return;
  }
  const std::to_chars_result r = std::to_chars(buf, buf + sizeof buf, c, std::chars_format::fixed, 6);
  out.append(buf, r.ptr);
}

// node.h:255-301
static int app(const Flat& f, int i, std::string& s, std::vector<int32_t>* be) {
  const Tok& t = f[i];
  if (be) (*be)[2 * i] = (int32_t)s.size();
  int nx = i + 1;
  switch (t.t) {
    case CONST: append_const(t.c, s); break;
    case VAR: s += 'x'; s += std::to_string(t.var); break;
    case PLUS: case MINUS: case MULTIPLY: case DIVIDE: case LT: case GT: case EQ: case NEQ: {
      static const char sym[] = {0, 0, 0, '+', '-', '*', '/', 0, 0, 0, 0, 0, 0, 0, '<', '<', '<', '<'};
      s += '(';
      nx = app(f, nx, s, be);
      s += sym[t.t];
      nx = app(f, nx, s, be);
      s += ')';
      break;
    }
    case SIN: s += "sin("; nx = app(f, nx, s, be); s += ')'; break;
    case COS: s += "cos("; nx = app(f, nx, s, be); s += ')'; break;
    case LN: s += "ln("; nx = app(f, nx, s, be); s += ')'; break;
    case EXP: s += "exp("; nx = app(f, nx, s, be); s += ')'; break;
    case SQRT: s += "sqrt("; nx = app(f, nx, s, be); s += ')'; break;
    case SQR: s += "(("; nx = app(f, nx, s, be); s += ")**2)"; break;
    case POW: s += "pow("; nx = app(f, nx, s, be); s += ','; nx = app(f, nx, s, be); s += ')'; break;
    case MIN: s += "MIN("; nx = app(f, nx, s, be); s += ", "; nx = app(f, nx, s, be); s += ')'; break;
    case MAX: s += "MAX("; nx = app(f, nx, s, be); s += ", "; nx = app(f, nx, s, be); s += ')'; break;
    default: s += "*****UNKNOWN*****"; break;
  }
  if (be) (*be)[2 * i + 1] = (int32_t)s.size();
  return nx;
}

void append_string(const Flat& f, int pos, std::string& out, std::vector<int32_t>* begin_end) {
  if (begin_end) begin_end->assign(2 * f.size(), 0);
  app(f, pos, out, begin_end);
}

std::string to_string(const Flat& f) {
  std::string s;
  s.reserve(f.size() * 6);
  app(f, 0, s, nullptr);
  return s;
}

// ---- exchange format
static const char* tok_name(uint8_t t) {
  switch (t) {
    case PLUS: return "+"; case MINUS: return "-"; case MULTIPLY: return "*"; case DIVIDE: return "/";
    case SIN: return "sin"; case COS: return "cos"; case LN: return "ln"; case EXP: return "exp";
    case SQRT: return "sqrt"; case SQR: return "sqr"; case POW: return "pow"; case LT: return "<";
    case GT: return ">"; case EQ: return "=="; case NEQ: return "!="; case MIN: return "min"; case MAX: return "max";
    default: return "?";
  }
}

std::string to_prefix(const Flat& f) {
  std::string s;
  char b[64];
  for (size_t i = 0; i < f.size(); ++i) {
    if (i) s += ' ';
    if (f[i].t == CONST) { snprintf(b, sizeof b, "c%a", f[i].c); s += b; }
    else if (f[i].t == VAR) { s += 'x'; s += std::to_string(f[i].var); }
    else s += tok_name(f[i].t);
  }
  return s;
}

Flat parse_prefix(const std::string& s) {
  Flat f;
  size_t p = 0;
  while (p < s.size()) {
    while (p < s.size() && s[p] == ' ') ++p;
    size_t q = p;
    while (q < s.size() && s[q] != ' ') ++q;
    if (q == p) break;
    const std::string t = s.substr(p, q - p);
    p = q;
    if (t.size() > 1 && t[0] == 'c' && t != "cos") { f.push_back(tok_const(strtod(t.c_str() + 1, nullptr))); continue; }
    if (t.size() > 1 && t[0] == 'x') { f.push_back(tok_var(atoi(t.c_str() + 1))); continue; }
    uint8_t ty = NONE;
    for (uint8_t k = PLUS; k <= MAX; ++k) if (t == tok_name(k)) { ty = k; break; }
    if (ty == NONE) throw std::runtime_error("bad prefix token: " + t);
    f.push_back(tok_op(ty));
  }
  std::vector<int32_t> sz;
  subtree_sizes(f, sz);
  if (f.empty() || sz[0] != (int)f.size()) throw std::runtime_error("malformed prefix tree");
  return f;
}

// ---- pool
int32_t Pool::mk(uint8_t t, int32_t l, int32_t r) {
  n.push_back(PNode{t, -1, 0.0, l, r});
  return (int32_t)n.size() - 1;
}
int32_t Pool::mkc(double c) {
  n.push_back(PNode{CONST, -1, c, -1, -1});
  return (int32_t)n.size() - 1;
}

int32_t Pool::load(const Flat& f, int pos, int* next) {
  const Tok& t = f[pos];
  const int32_t me = (int32_t)n.size();
  n.push_back(PNode{t.t, t.var, t.c, -1, -1});
  int p = pos + 1;
  const int a = arity(t.t);
  if (a >= 1) { int32_t l = load(f, p, &p); n[me].l = l; }
  if (a >= 2) { int32_t r = load(f, p, &p); n[me].r = r; }
  if (next) *next = p;
  return me;
}

int32_t Pool::clone(int32_t i) {
  const PNode src = n[i];
  const int32_t me = (int32_t)n.size();
  n.push_back(PNode{src.t, src.var, src.c, -1, -1});
  if (src.l >= 0) { int32_t l = clone(src.l); n[me].l = l; }
  if (src.r >= 0) { int32_t r = clone(src.r); n[me].r = r; }
  return me;
}

void Pool::store(int32_t i, Flat& out) const {
  const PNode& p = n[i];
  out.push_back(Tok{p.t, p.var, p.c});
  const int a = arity(p.t);
  if (a >= 1) store(p.l, out);
  if (a >= 2) store(p.r, out);
}

int Pool::size(int32_t i) const {   // node.h:311-322
  const int a = arity(n[i].t);
  if (a == 0) return 1;
  if (a == 1) return 1 + size(n[i].l);
  return 1 + size(n[i].l) + size(n[i].r);
}

// node::update_with (node.h:130-139): self takes src's type, values and children
void Pool::update_with(int32_t self, int32_t src) {
  const PNode s = n[src];
  PNode& d = n[self];
  d.t = s.t; d.c = s.c; d.var = s.var; d.l = s.l; d.r = s.r;
}

// node.cpp:152-296
void Pool::simplify(int32_t i) {
  const int a = arity(n[i].t);
  if (a == 0) return;
  if (a == 1) { simplify(n[i].l); return; }
  simplify(n[i].l);
  simplify(n[i].r);
  const uint8_t t = n[i].t;
  const int32_t L = n[i].l, R = n[i].r;
  if (n[L].t == CONST && n[R].t == CONST) {
    const double x = n[L].c, y = n[R].c;
    double v;
    switch (t) {
      case PLUS: v = x + y; break;
      case MINUS: v = x - y; break;
      case MULTIPLY: v = x * y; break;
      case DIVIDE: v = x / y; break;
      case POW: v = std::pow(x, y); break;
      case LT: v = x < y ? 1 : 0; break;
      case GT: v = x > y ? 1 : 0; break;
      case EQ: v = x == y ? 1 : 0; break;
      case NEQ: v = x != y ? 1 : 0; break;
      case MIN: v = x < y ? x : y; break;
      case MAX: v = x > y ? x : y; break;
      default: throw std::runtime_error("Simplification is not supported for this binary operator!");
    }
    n[i].c = v; n[i].t = CONST; n[i].l = -1; n[i].r = -1;
    return;
  }
  if (n[L].t == CONST) {
    if (t == PLUS || t == MINUS) {
      if (value_zero(n[L].c)) { update_with(i, R); return; }
      if (n[R].t == PLUS || n[R].t == MINUS) {
        const int32_t RL = n[R].l, RR = n[R].r;
        if (n[RL].t == CONST) {
          if (t == PLUS) n[L].c += n[RL].c; else n[L].c -= n[RL].c;
          update_with(R, RR);
        } else if (n[RR].t == CONST) {
          if (t == PLUS) { if (n[R].t == PLUS) n[L].c += n[RR].c; else n[L].c -= n[RR].c; }
          else { if (n[R].t == PLUS) n[L].c -= n[RR].c; else n[L].c += n[RR].c; }
          update_with(R, RL);
        }
      }
    } else if (t == MULTIPLY) {
      if (value_zero(n[L].c)) { n[i].t = CONST; n[i].c = 0.0; n[i].l = n[i].r = -1; }
      else if (value_one(n[L].c)) update_with(i, R);
      else if (n[R].t == MULTIPLY) {
        const int32_t RL = n[R].l, RR = n[R].r;
        if (n[RL].t == CONST) { n[L].c *= n[RL].c; update_with(R, RR); }
        else if (n[RR].t == CONST) { n[L].c *= n[RR].c; update_with(R, RL); }
      }
    } else if (t == DIVIDE && value_zero(n[L].c)) {
      n[i].t = CONST; n[i].c = 0.0; n[i].l = n[i].r = -1;
    }
    return;
  }
  if (n[R].t == CONST) {
    if (t == PLUS || t == MINUS) {
      if (value_zero(n[R].c)) { update_with(i, L); return; }
      if (n[L].t == PLUS || n[L].t == MINUS) {
        const int32_t LL = n[L].l, LR = n[L].r;
        if (n[LL].t == CONST) {
          if (t == PLUS) n[LL].c += n[R].c; else n[LL].c -= n[R].c;
          n[R].c = 0;
        } else if (n[LR].t == CONST) {
          if (n[L].t == PLUS) n[R].c += n[LR].c; else n[R].c -= n[LR].c;
          n[LR].c = 0;
        }
      }
    } else if (t == MULTIPLY) {
      if (value_one(n[R].c)) update_with(i, L);
      else if (value_zero(n[R].c)) { n[i].t = CONST; n[i].c = 0.0; n[i].l = n[i].r = -1; }
      else if (n[L].t == MULTIPLY) {
        const int32_t LL = n[L].l, LR = n[L].r;
        if (n[LL].t == CONST) { n[R].c *= n[LL].c; update_with(L, LR); }
        else if (n[LR].t == CONST) { n[R].c *= n[LR].c; update_with(L, LL); }
      }
    }
  }
}

// node.cpp:312-327
void Pool::normalize_factor_constants(int32_t i, bool inside) {
  const uint8_t t = n[i].t;
  if (t == CONST) n[i].c = 1;
  else if (!inside && (t == PLUS || t == MINUS)) {
    normalize_factor_constants(n[i].l, false);
    normalize_factor_constants(n[i].r, false);
  } else if (!inside) {
    if (t == MULTIPLY) {
      normalize_factor_constants(n[i].l, true);
      normalize_factor_constants(n[i].r, true);
    } else if (t == DIVIDE) {
      normalize_factor_constants(n[i].r, true);
    }
  }
}

// node.cpp:329-338
int32_t Pool::binomial_mult(int32_t left, int32_t right) {
  auto mul = [&](int32_t a, int32_t b) {   // MULTIPLY(copy a, copy b)
    const int32_t ca = clone(a);
    const int32_t cb = clone(b);
    return mk(MULTIPLY, ca, cb);
  };
  const int32_t f1 = mul(n[left].l, n[right].l);
  const int32_t f2 = mul(n[left].l, n[right].r);
  const int32_t f3 = mul(n[left].r, n[right].l);
  const int32_t f4 = mul(n[left].r, n[right].r);
  const int32_t nl = mk(n[right].t, f1, f2);
  const int32_t nr = mk(n[right].t, f3, f4);
  return mk(n[left].t, nl, nr);
}

// node.cpp:340-374, with the one-sided distributions exactly as written there
void Pool::expand(int32_t i) {
  const int a = arity(n[i].t);
  if (a == 0) return;
  if (a == 1) { expand(n[i].l); return; }
  expand(n[i].l);
  expand(n[i].r);
  if (n[i].t != MULTIPLY) return;
  const int32_t L = n[i].l, R = n[i].r;
  const bool lsum = n[L].t == PLUS || n[L].t == MINUS;
  const bool rsum = n[R].t == PLUS || n[R].t == MINUS;
  if (lsum) {
    if (rsum) {
      const int32_t b = binomial_mult(L, R);
      update_with(i, b);
    } else {
      // left = t1*t3 ; right = (new left)->right * t3 ; type = new left's type
      const int32_t t1 = clone(n[L].l);
      const int32_t nl = mk(MULTIPLY, t1, -1);
      { const int32_t c3 = clone(R); n[nl].r = c3; }
      n[i].l = nl;
      const int32_t a1 = clone(n[nl].r);
      const int32_t nr = mk(MULTIPLY, a1, -1);
      { const int32_t c3 = clone(R); n[nr].r = c3; }
      n[i].r = nr;
      n[i].t = n[nl].t;
    }
  } else if (rsum) {
    // left = t1*t2 ; right = (new left) * t3 ; type = new right's type
    const int32_t t1 = clone(L);
    const int32_t nl = mk(MULTIPLY, t1, -1);
    { const int32_t c2 = clone(n[R].l); n[nl].r = c2; }
    n[i].l = nl;
    const int32_t a1 = clone(nl);
    const int32_t nr = mk(MULTIPLY, a1, -1);
    { const int32_t c3 = clone(n[R].r); n[nr].r = c3; }
    n[i].r = nr;
    n[i].t = n[nr].t;
  }
}

// node.cpp:140-147
void Pool::extract_factors(int32_t i, std::vector<int32_t>& out) const {
  const uint8_t t = n[i].t;
  if (t == PLUS || t == MINUS) {
    extract_factors(n[i].l, out);
    extract_factors(n[i].r, out);
  } else if (t != CONST) {
    out.push_back(i);
  }
}

void rewrite_rounds(Flat& f, bool normalize, Pool& pool) {
  pool.clear();
  const int32_t root = pool.load(f);
  int it_max = 5;
  while (it_max > 0) {
    const int sz = pool.size(root);
    pool.expand(root);
    if (normalize) pool.normalize_factor_constants(root, false);
    pool.simplify(root);
    if (pool.size(root) == sz) break;
    --it_max;
  }
  f.clear();
  pool.store(root, f);
}

void ols_factors(const Flat& cand, std::vector<Flat>& factors, Pool& pool) {
  pool.clear();
  const int32_t root = pool.load(cand);
  pool.expand(root);
  pool.simplify(root);
  std::vector<int32_t> all;
  pool.extract_factors(root, all);
  factors.clear();
  auto take = [&](int32_t i) { factors.emplace_back(); pool.store(i, factors.back()); };
  for (int32_t f : all) {
    const PNode& p = pool.n[f];
    if (p.t == CONST) continue;
    const int a = arity(p.t);
    if (a == 2 && pool.n[p.l].t == CONST && pool.n[p.r].t == CONST) continue;
    if (p.t == MULTIPLY || p.t == PLUS || p.t == MINUS) {
      if (pool.n[p.l].t == CONST) { take(p.r); continue; }
      if (pool.n[p.r].t == CONST) { take(p.l); continue; }
    }
    if (p.t == DIVIDE && pool.n[p.r].t == CONST) { take(p.l); continue; }
    take(f);
  }
}

void ols_tree(const std::vector<Flat>& factors, const double* coefs, Flat& out) {
  out.clear();
  const int nf = (int)factors.size();
  int kept = 0;
  for (int i = 0; i <= nf; ++i) if (!value_zero(coefs[i])) ++kept;
  if (kept == 0) { out.push_back(tok_const(0.0)); return; }
  for (int i = 1; i < kept; ++i) out.push_back(tok_op(PLUS));
  for (int i = 0; i < nf; ++i) {
    const double c = coefs[i];
    if (value_zero(c)) continue;
    if (!value_one(c)) { out.push_back(tok_op(MULTIPLY)); out.push_back(tok_const(c)); }
    out.insert(out.end(), factors[i].begin(), factors[i].end());
  }
  if (!value_zero(coefs[nf])) out.push_back(tok_const(coefs[nf] * 1.0));
}

int ols_tree_size(const std::vector<int32_t>& fs, const double* coefs) {
  const int nf = (int)fs.size();
  int kept = 0, sz = 0;
  for (int i = 0; i < nf; ++i) {
    if (value_zero(coefs[i])) continue;
    ++kept;
    sz += value_one(coefs[i]) ? fs[i] : fs[i] + 2;
  }
  if (!value_zero(coefs[nf])) { ++kept; sz += 1; }
  if (kept == 0) return 1;
  return sz + kept - 1;
}

// ---- bytecode
namespace {
struct Emitter {
  const Flat& f;
  const std::vector<int32_t>& sz;
  std::vector<int32_t> need;
  std::vector<uint32_t>& code;
  std::vector<double>& consts;
  int sp = 0, depth = 0;

  static bool leaf(uint8_t t) { return t == CONST || t == VAR; }
  int right(int i) const { return i + 1 + sz[i + 1]; }
  void emit(uint32_t op, uint32_t arg = 0) { code.push_back((arg << 8) | op); }
  uint32_t kidx(double c) { consts.push_back(c); return (uint32_t)consts.size() - 1; }

  static uint32_t bop(uint8_t t) {
    switch (t) {
      case PLUS: return RR_BOP_ADD; case MINUS: return RR_BOP_SUB; case MULTIPLY: return RR_BOP_MUL;
      case DIVIDE: return RR_BOP_DIV; case LT: return RR_BOP_LT; case GT: return RR_BOP_GT;
      case EQ: return RR_BOP_EQ; case NEQ: return RR_BOP_NEQ; case MIN: return RR_BOP_MIN; case MAX: return RR_BOP_MAX;
      default: throw std::runtime_error("operator has no device form (POW is never generated)");
    }
  }
  static uint32_t uop(uint8_t t) {
    switch (t) {
      case SIN: return RR_OP_SIN; case COS: return RR_OP_COS; case LN: return RR_OP_LN;
      case EXP: return RR_OP_EXP; case SQRT: return RR_OP_SQRT; default: return RR_OP_SQR;
    }
  }
  void compute_need() {
    const int m = (int)f.size();
    need.assign(m, 0);
    for (int i = m - 1; i >= 0; --i) {
      const int a = arity(f[i].t);
      if (a == 0) need[i] = 0;
      else if (a == 1) need[i] = need[i + 1];
      else {
        const int l = i + 1, r = right(i);
        if (leaf(f[r].t)) need[i] = need[l];
        else if (leaf(f[l].t)) need[i] = need[r];
        else need[i] = need[l] == need[r] ? need[l] + 1 : std::max(need[l], need[r]);
      }
    }
  }
  void gen(int i) {
    const Tok& t = f[i];
    const int a = arity(t.t);
    if (t.t == VAR) { emit(RR_OP_LDV, (uint32_t)t.var); return; }
    if (t.t == CONST) { emit(RR_OP_LDC, kidx(t.c)); return; }
    if (a == 1) { gen(i + 1); emit(uop(t.t)); return; }
    const int l = i + 1, r = right(i);
    const uint32_t base = RR_OP_BIN + 8 * bop(t.t);
    if (leaf(f[r].t)) {
      gen(l);
      if (f[r].t == VAR) emit(base + RR_M_V, (uint32_t)f[r].var); else emit(base + RR_M_C, kidx(f[r].c));
    } else if (leaf(f[l].t)) {
      gen(r);
      if (f[l].t == VAR) emit(base + RR_M_VR, (uint32_t)f[l].var); else emit(base + RR_M_CR, kidx(f[l].c));
    } else if (need[l] >= need[r]) {
      gen(l); emit(RR_OP_PUSH); depth = std::max(depth, ++sp);
      gen(r); emit(base + RR_M_S); --sp;
    } else {
      gen(r); emit(RR_OP_PUSH); depth = std::max(depth, ++sp);
      gen(l); emit(base + RR_M_SR); --sp;
    }
  }
};
}  // namespace

int compile(const Flat& f, std::vector<uint32_t>& code, std::vector<double>& consts) {
  std::vector<int32_t> sz;
  subtree_sizes(f, sz);
  Emitter e{f, sz, {}, code, consts};
  e.compute_need();
  e.gen(0);
  if (e.depth > RR_MAX_STACK) throw std::runtime_error("expression needs more operand-stack slots than RR_MAX_STACK");
  return e.depth;
}

}  // namespace rrh
