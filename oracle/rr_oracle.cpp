// rr_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU FP64 restatement of the RILS-ROLS candidate-fitness path and of the
// search around it, used as the checker for the GPU implementation.  It is
// loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
// leg, and by nothing on the product path.
//
// It restates (file:line in the reference):
//   node type, arity, symmetry, allowed-left rule .... rils_rols_cpp/node.h:16-68, node.cpp:97-118
//   to_string / size .................................. node.h:255-322
//   evaluate_all / evaluate_inner ..................... node.cpp:5-95
//   simplify .......................................... node.cpp:152-296
//   normalize_factor_constants ........................ node.cpp:312-327
//   expand / binomial_mult ............................ node.cpp:329-374
//   extract_non_constant_factors ...................... node.cpp:140-147
//   get_all_subtrees .................................. node.h:351-377
//   R2 / RMSE ......................................... rils_rols_cpp.cpp:40-49
//   setup_nodes, generators, all_candidates ........... rils_rols_cpp.cpp:121-443
//   tune_constants / fitness .......................... rils_rols_cpp.cpp:445-541
//   fitness_value, compare_fitness, pareto ............ rils_rols_cpp.cpp:543-599
//   local_search ...................................... rils_rols_cpp.cpp:601-643
//   finished / check_skip ............................. rils_rols_cpp.cpp:659-672
//   get_x / get_y / predict ........................... rils_rols_cpp.cpp:675-750
//   relevant_features / fit_inner ..................... rils_rols_cpp.cpp:753-859
//   ColPivHouseholderQR::computeInPlace / _solve_impl . eigen/Eigen/src/QR/ColPivHouseholderQR.h:482-607
//   makeHouseholder / applyHouseholderOnTheLeft ....... eigen/Eigen/src/Householder/Householder.h:43-135
//   HouseholderSequence::applyThisOnTheLeft ........... eigen/Eigen/src/Householder/HouseholderSequence.h:369-413
//
// Arithmetic conventions (the reference's are partly unreadable: Eigen's
// reduction kernels and vectorised transcendentals live in Eigen/Core, which
// the vendored copy lacks):
//   * every sum over samples uses one canonical order: 1024 partial sums,
//     row i added to partial (i mod 1024) in increasing i, each group of 64
//     partials combined by an xor-butterfly (pairs at distance 32, 16, 8, 4,
//     2, 1), the 16 group sums added in order; rrm::canon_sum (rr_math.h).
//     Every sum of the column-pivoted QR (norms, tail sums, Householder dot
//     products) and the fitness's sum of squared residuals use the chunked
//     order of rrm::chunk_sum instead (chunks of 4096 rows, 256 partials
//     each, the chunk sums added in order): the GPU does each pass over a
//     candidate's rows as independent chunks;
//   * maxCoeff() is a left-to-right scan with strict '>' (first index wins);
//   * sin/cos/exp/log come from rr_math.h; + - * / sqrt are IEEE;
//   * no floating-point contraction (built with -ffp-contract=off);
//   * a NaN coefficient leaving the solver is replaced by the x86 default
//     NaN, so model strings print it as "-nan" whatever produced it.
//
// What pins each part (the reference holds no golden vectors for this path
// and cannot be built or imported here; tests/test_oracle.py unless noted):
//   sin/cos/exp/log ......... < 1 ulp of mpmath over wide ranges; special
//                             values (test_math_*)
//   canonical sums .......... exact on integer data, within the fsum error
//                             bound otherwise; the ORDER itself is this
//                             project's convention (Eigen's is unreadable):
//                             last-bit agreement with the reference is
//                             parity unpinned
//   ColPivHouseholderQR ..... Eigen's order on the data (more than 9
//                             columns, and the direct mode): numpy lstsq on
//                             full-rank designs, exact integer coefficients,
//                             duplicated / constant / zero columns, n < k,
//                             pivot order of LAPACK geqp3 (test_qr_*)
//   chunked QR (<= 9 cols) .. lstsq and exact integer fits over 1 to 49
//                             chunks (1 to 4 groups) in every column bucket;
//                             the same rank, zero pattern and coefficients as
//                             Eigen's order on designs whose rank cut is clear
//                             (columns living in one chunk or one group,
//                             zero columns, wide designs); NaN / Inf outcomes.
//                             Exactly collinear columns leave a residual at
//                             the size of the cut itself, decided by rounding
//                             in every order, Eigen's included
//                             (test_exact_collinearity_rank_is_an_order_decision):
//                             agreement with the reference there is parity
//                             unpinned
//   score / R2 / RMSE ....... closed forms: exact fits, formulas, NaN sentinel,
//                             Inf passthrough
//   rewrites, to_string ..... known answers derived by hand from node.cpp,
//                             faithful bugs included
//   all_candidates .......... known answers: the main loop's perturbations of
//                             x0, the local-search lists of (x0+2) and
//                             ln(sin(x0)) (tests/test_search_pins.py)
//   acceptance primitives ... logic tables (test_acceptance_primitives_*)
//   glibc rand sequence ..... the same library call (test_crand_is_glibc_rand
//                             in tests/test_host_engine.py)
//   local_search ............ a search hand-traced on a tiny data set:
//                             acceptance, pareto, budget cuts, the extra fit
//                             call (tests/test_search_pins.py)
//   fit_inner ............... determinism and fit-call budget accounting;
//                             exact recovery of the generating formula on seven
//                             of the reference's own noise-free datasets, where
//                             the reference reports R2 = 1 (tests/test_recovery.py);
//                             the move sequence of longer runs is checked for
//                             agreement with the product only
#define RR_HD
#include "../rils-rols_amd/csrc/rr_math.h"

#include <algorithm>
#include <atomic>
#include <thread>
#include <chrono>
#include <climits>
#include <limits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <tuple>
#include <unordered_set>
#include <vector>

namespace orc {

static const double EPS = 1e-12;       // node.h:13-14: pow(10, -12)
static const double PI_ = 3.14159265358979323846;  // node.h:11

enum class T : int {
  NONE, CONST, VAR, PLUS, MINUS, MULTIPLY, DIVIDE, SIN, COS, LN, EXP, SQRT, SQR, POW,
  LT, GT, EQ, NEQ, MIN, MAX
};

static int arity(T t) {
  switch (t) {
    case T::CONST: case T::VAR: return 0;
    case T::SIN: case T::COS: case T::LN: case T::EXP: case T::SQRT: case T::SQR: return 1;
    default: return 2;
  }
}
static bool symmetric(T t) { return !(t == T::MINUS || t == T::DIVIDE || t == T::POW); }

struct Node;
using P = std::unique_ptr<Node>;

struct Node {
  T type = T::NONE;
  int var = -1;
  double c = 0.0;
  P l, r;
  Node() = default;
  explicit Node(T t) : type(t) {}
  int ar() const { return arity(type); }
  P clone() const {
    P n(new Node(type));
    n->var = var; n->c = c;
    if (l) n->l = l->clone();
    if (r) n->r = r->clone();
    return n;
  }
};

static P mk(T t) { return P(new Node(t)); }
static P mkc(double v) { P n = mk(T::CONST); n->c = v; return n; }
static P mkv(int i) { P n = mk(T::VAR); n->var = i; return n; }
static P mk2(T t, P a, P b) { P n = mk(t); n->l = std::move(a); n->r = std::move(b); return n; }
static P mk1(T t, P a) { P n = mk(t); n->l = std::move(a); return n; }

static bool value_zero(double v) { return std::fabs(v) < EPS; }
static bool value_one(double v) { return std::fabs(v - 1) < EPS; }

// node.cpp:97-118
static bool allowed_left(T parent, T child) {
  switch (parent) {
    case T::EXP: case T::LN: return !(child == T::EXP || child == T::LN);
    case T::POW: return child != T::POW;
    case T::COS: case T::SIN: return !(child == T::COS || child == T::SIN);
    default: return true;
  }
}

// node.h:311-322
static int size(const Node& n) {
  const int a = n.ar();
  if (a == 0) return 1;
  if (a == 1) return 1 + size(*n.l);
  return 1 + size(*n.l) + size(*n.r);
}

// std::to_string(double) is printf("%f"); (int)round(c) of an out-of-range
// value is what x86 cvttsd2si gives: INT_MIN.
static std::string const_str(double c) {
  const double rc = std::round(c);
  if (std::fabs(rc - c) < EPS) {
    int iv = (rc >= -2147483648.0 && rc < 2147483648.0) ? (int)rc : INT_MIN;
    return std::to_string(iv);
  }
  return std::to_string(c);
}

// node.h:255-301 (including '<' printed for >, ==, !=)
static void to_str(const Node& n, std::string& s) {
  switch (n.type) {
    case T::CONST: s += const_str(n.c); return;
    case T::VAR: s += "x"; s += std::to_string(n.var); return;
    case T::PLUS: s += "("; to_str(*n.l, s); s += "+"; to_str(*n.r, s); s += ")"; return;
    case T::MINUS: s += "("; to_str(*n.l, s); s += "-"; to_str(*n.r, s); s += ")"; return;
    case T::MULTIPLY: s += "("; to_str(*n.l, s); s += "*"; to_str(*n.r, s); s += ")"; return;
    case T::DIVIDE: s += "("; to_str(*n.l, s); s += "/"; to_str(*n.r, s); s += ")"; return;
    case T::SIN: s += "sin("; to_str(*n.l, s); s += ")"; return;
    case T::COS: s += "cos("; to_str(*n.l, s); s += ")"; return;
    case T::LN: s += "ln("; to_str(*n.l, s); s += ")"; return;
    case T::EXP: s += "exp("; to_str(*n.l, s); s += ")"; return;
    case T::SQRT: s += "sqrt("; to_str(*n.l, s); s += ")"; return;
    case T::SQR: s += "(("; to_str(*n.l, s); s += ")**2)"; return;
    case T::POW: s += "pow("; to_str(*n.l, s); s += ","; to_str(*n.r, s); s += ")"; return;
    case T::LT: case T::GT: case T::EQ: case T::NEQ:
      s += "("; to_str(*n.l, s); s += "<"; to_str(*n.r, s); s += ")"; return;
    case T::MIN: s += "MIN("; to_str(*n.l, s); s += ", "; to_str(*n.r, s); s += ")"; return;
    case T::MAX: s += "MAX("; to_str(*n.l, s); s += ", "; to_str(*n.r, s); s += ")"; return;
    default: s += "*****UNKNOWN*****"; return;
  }
}
static std::string to_string(const Node& n) { std::string s; to_str(n, s); return s; }

// ------------------------------------------------------------ prefix codec
// Exchange format: space-separated prefix tokens; constants as "c<%a>".
static const char* tok_name(T t) {
  switch (t) {
    case T::PLUS: return "+"; case T::MINUS: return "-"; case T::MULTIPLY: return "*";
    case T::DIVIDE: return "/"; case T::SIN: return "sin"; case T::COS: return "cos";
    case T::LN: return "ln"; case T::EXP: return "exp"; case T::SQRT: return "sqrt";
    case T::SQR: return "sqr"; case T::POW: return "pow"; case T::LT: return "<";
    case T::GT: return ">"; case T::EQ: return "=="; case T::NEQ: return "!=";
    case T::MIN: return "min"; case T::MAX: return "max"; default: return "?";
  }
}
static void to_prefix(const Node& n, std::string& s) {
  if (!s.empty()) s += ' ';
  if (n.type == T::CONST) { char b[64]; snprintf(b, sizeof b, "c%a", n.c); s += b; return; }
  if (n.type == T::VAR) { s += "x"; s += std::to_string(n.var); return; }
  s += tok_name(n.type);
  if (n.ar() >= 1) to_prefix(*n.l, s);
  if (n.ar() >= 2) to_prefix(*n.r, s);
}
static std::string prefix(const Node& n) { std::string s; to_prefix(n, s); return s; }

static P parse_rec(const std::vector<std::string>& tk, size_t& i) {
  if (i >= tk.size()) throw std::runtime_error("prefix: truncated");
  const std::string& t = tk[i++];
  if (t[0] == 'c' && t.size() > 1 && t != "cos") return mkc(strtod(t.c_str() + 1, nullptr));
  if (t[0] == 'x' && t.size() > 1 && t != "exp") return mkv(atoi(t.c_str() + 1));
  static const T all[] = {T::PLUS, T::MINUS, T::MULTIPLY, T::DIVIDE, T::SIN, T::COS, T::LN, T::EXP,
                          T::SQRT, T::SQR, T::POW, T::LT, T::GT, T::EQ, T::NEQ, T::MIN, T::MAX};
  for (T ty : all) {
    if (t == tok_name(ty)) {
      P n = mk(ty);
      if (n->ar() >= 1) n->l = parse_rec(tk, i);
      if (n->ar() >= 2) n->r = parse_rec(tk, i);
      return n;
    }
  }
  throw std::runtime_error("prefix: bad token " + t);
}
static P parse(const std::string& s) {
  std::vector<std::string> tk;
  size_t p = 0;
  while (p < s.size()) {
    while (p < s.size() && s[p] == ' ') ++p;
    size_t q = p;
    while (q < s.size() && s[q] != ' ') ++q;
    if (q > p) tk.push_back(s.substr(p, q - p));
    p = q;
  }
  size_t i = 0;
  P n = parse_rec(tk, i);
  if (i != tk.size()) throw std::runtime_error("prefix: trailing tokens");
  return n;
}

// ------------------------------------------------------------ rewrites
// update_with (node.h:130-139): this takes src's content (src is a descendant or new)
static void update_with(Node& self, P src) {
  self.type = src->type;
  self.c = src->c;
  self.var = src->var;
  P nl = std::move(src->l), nr = std::move(src->r);
  self.l = std::move(nl);
  self.r = std::move(nr);
}
static void set_const_value(Node& self, double v) {  // node.h:222-228
  self.type = T::CONST; self.c = v; self.l.reset(); self.r.reset();
}

// node.cpp:152-296
static void simplify(Node& n) {
  const int a = n.ar();
  if (a == 0) return;
  if (a == 1) { simplify(*n.l); return; }
  simplify(*n.l);
  simplify(*n.r);
  Node& L = *n.l;
  Node& R = *n.r;
  if (L.type == T::CONST && R.type == T::CONST) {
    const double x = L.c, y = R.c;
    double v;
    switch (n.type) {
      case T::PLUS: v = x + y; break;
      case T::MINUS: v = x - y; break;
      case T::MULTIPLY: v = x * y; break;
      case T::DIVIDE: v = x / y; break;
      case T::POW: v = std::pow(x, y); break;
      case T::LT: v = x < y ? 1 : 0; break;
      case T::GT: v = x > y ? 1 : 0; break;
      case T::EQ: v = x == y ? 1 : 0; break;
      case T::NEQ: v = x != y ? 1 : 0; break;
      case T::MIN: v = x < y ? x : y; break;
      case T::MAX: v = x > y ? x : y; break;
      default: throw std::runtime_error("Simplification is not supported for this binary operator!");
    }
    n.c = v; n.type = T::CONST; n.l.reset(); n.r.reset();
  } else if (L.type == T::CONST) {
    if (n.type == T::PLUS || n.type == T::MINUS) {
      if (value_zero(L.c)) {
        update_with(n, std::move(n.r));
      } else if (R.type == T::PLUS || R.type == T::MINUS) {
        if (R.l->type == T::CONST) {
          if (n.type == T::PLUS) L.c += R.l->c; else L.c -= R.l->c;
          update_with(R, std::move(R.r));
        } else if (R.r->type == T::CONST) {
          if (n.type == T::PLUS) {
            if (R.type == T::PLUS) L.c += R.r->c; else L.c -= R.r->c;
          } else {
            if (R.type == T::PLUS) L.c -= R.r->c; else L.c += R.r->c;
          }
          update_with(R, std::move(R.l));
        }
      }
    } else if (n.type == T::MULTIPLY) {
      if (value_zero(L.c)) set_const_value(n, 0.0);
      else if (value_one(L.c)) update_with(n, std::move(n.r));
      else if (R.type == T::MULTIPLY) {
        if (R.l->type == T::CONST) { L.c *= R.l->c; update_with(R, std::move(R.r)); }
        else if (R.r->type == T::CONST) { L.c *= R.r->c; update_with(R, std::move(R.l)); }
      }
    } else if (n.type == T::DIVIDE && value_zero(L.c)) {
      set_const_value(n, 0.0);
    }
  } else if (R.type == T::CONST) {
    if (n.type == T::PLUS || n.type == T::MINUS) {
      if (value_zero(R.c)) {
        update_with(n, std::move(n.l));
      } else if (L.type == T::PLUS || L.type == T::MINUS) {
        if (L.l->type == T::CONST) {
          if (n.type == T::PLUS) L.l->c += R.c; else L.l->c -= R.c;
          R.c = 0;
        } else if (L.r->type == T::CONST) {
          if (L.type == T::PLUS) R.c += L.r->c; else R.c -= L.r->c;
          L.r->c = 0;
        }
      }
    } else if (n.type == T::MULTIPLY) {
      if (value_one(R.c)) update_with(n, std::move(n.l));
      else if (value_zero(R.c)) set_const_value(n, 0.0);
      else if (L.type == T::MULTIPLY) {
        if (L.l->type == T::CONST) { R.c *= L.l->c; update_with(L, std::move(L.r)); }
        else if (L.r->type == T::CONST) { R.c *= L.r->c; update_with(L, std::move(L.l)); }
      }
    }
  }
}

// node.cpp:312-327
static void normalize_factor_constants(Node& n, T parent, bool inside) {
  (void)parent;
  if (n.type == T::CONST) n.c = 1;
  else if (!inside && (n.type == T::PLUS || n.type == T::MINUS)) {
    normalize_factor_constants(*n.l, n.type, false);
    normalize_factor_constants(*n.r, n.type, false);
  } else if (!inside) {
    if (n.type == T::MULTIPLY) {
      normalize_factor_constants(*n.l, n.type, true);
      normalize_factor_constants(*n.r, n.type, true);
    } else if (n.type == T::DIVIDE) {
      normalize_factor_constants(*n.r, n.type, true);
    }
  }
}

// node.cpp:329-338
static P binomial_mult(const Node& left, const Node& right) {
  P f1 = mk2(T::MULTIPLY, left.l->clone(), right.l->clone());
  P f2 = mk2(T::MULTIPLY, left.l->clone(), right.r->clone());
  P f3 = mk2(T::MULTIPLY, left.r->clone(), right.l->clone());
  P f4 = mk2(T::MULTIPLY, left.r->clone(), right.r->clone());
  P nl = mk2(right.type, std::move(f1), std::move(f2));
  P nr = mk2(right.type, std::move(f3), std::move(f4));
  return mk2(left.type, std::move(nl), std::move(nr));
}

// node.cpp:340-374, including the two one-sided distributions as written
// there: (t1 +- t2)*t3 becomes (t1*t3)*(t3*t3) and t1*(t2 +- t3) becomes
// (t1*t2)*((t1*t2)*t3).
static void expand(Node& n) {
  const int a = n.ar();
  if (a == 0) return;
  if (a == 1) { expand(*n.l); return; }
  expand(*n.l);
  expand(*n.r);
  if (n.type != T::MULTIPLY) return;
  const bool lsum = n.l->type == T::PLUS || n.l->type == T::MINUS;
  const bool rsum = n.r->type == T::PLUS || n.r->type == T::MINUS;
  if (lsum) {
    if (rsum) {
      update_with(n, binomial_mult(*n.l, *n.r));
    } else {
      P newl = mk2(T::MULTIPLY, n.l->l->clone(), n.r->clone());   // left = t1*t3
      n.l = std::move(newl);
      P newr = mk2(T::MULTIPLY, n.l->r->clone(), n.r->clone());   // right = (new left's right)*t3
      n.r = std::move(newr);
      n.type = n.l->type;
    }
  } else if (rsum) {
    P newl = mk2(T::MULTIPLY, n.l->clone(), n.r->l->clone());     // left = t1*t2
    n.l = std::move(newl);
    P newr = mk2(T::MULTIPLY, n.l->clone(), n.r->r->clone());     // right = (new left)*t3
    n.r = std::move(newr);
    n.type = n.r->type;
  }
}

// node.cpp:140-147
static void extract_factors(Node& n, std::vector<Node*>& out) {
  if (n.type == T::PLUS || n.type == T::MINUS) {
    extract_factors(*n.l, out);
    extract_factors(*n.r, out);
  } else if (n.type != T::CONST) {
    out.push_back(&n);
  }
}

// node.h:351-377 (each entry an independent tree)
static void all_subtrees(const Node& root, std::vector<P>& out) {
  size_t pos = out.size();
  out.push_back(root.clone());
  while (pos < out.size()) {
    const Node* cur = out[pos].get();
    if (cur->l) out.push_back(cur->l->clone());
    if (cur->r) out.push_back(cur->r->clone());
    ++pos;
  }
}

// ------------------------------------------------------------ data, sums
struct Data {
  int64_t n = 0;
  int d = 0;
  std::vector<std::vector<double>> X;  // feature-major
  std::vector<double> y;
  double ybar = 0, sst = 0;
};

// the canonical order of sums over samples (rr_math.h)
template <class F>
static double canon_sum(int64_t lo, int64_t hi, F f) {
  return rrm::canon_sum(lo, hi, f);
}

static void finish_data(Data& D) {
  const std::vector<double>& y = D.y;
  D.ybar = canon_sum(0, D.n, [&](int64_t i) { return y[i]; }) / (double)D.n;
  const double yb = D.ybar;
  D.sst = canon_sum(0, D.n, [&](int64_t i) { return (y[i] - yb) * (y[i] - yb); });
}

// node.cpp:5-95
static std::vector<double> evaluate(const Node& nd, const Data& D) {
  const int64_t n = D.n;
  std::vector<double> out(n);
  switch (nd.type) {
    case T::CONST: std::fill(out.begin(), out.end(), nd.c); return out;
    case T::VAR:
      if (nd.var < 0 || nd.var >= D.d) throw std::runtime_error("variable index out of range");
      return D.X[nd.var];
    default: break;
  }
  const std::vector<double> a = evaluate(*nd.l, D);
  if (nd.ar() == 1) {
    switch (nd.type) {
      case T::SIN: for (int64_t i = 0; i < n; ++i) out[i] = rrm::sin(a[i]); break;
      case T::COS: for (int64_t i = 0; i < n; ++i) out[i] = rrm::cos(a[i]); break;
      case T::LN: for (int64_t i = 0; i < n; ++i) out[i] = rrm::log(a[i]); break;
      case T::EXP: for (int64_t i = 0; i < n; ++i) out[i] = rrm::exp(a[i]); break;
      case T::SQRT: for (int64_t i = 0; i < n; ++i) out[i] = rrm::sqrt(a[i]); break;
      case T::SQR: for (int64_t i = 0; i < n; ++i) out[i] = a[i] * a[i]; break;
      default: throw std::runtime_error("ERROR: Unrecognized operation.");
    }
    return out;
  }
  const std::vector<double> b = evaluate(*nd.r, D);
  switch (nd.type) {
    case T::PLUS: for (int64_t i = 0; i < n; ++i) out[i] = a[i] + b[i]; break;
    case T::MINUS: for (int64_t i = 0; i < n; ++i) out[i] = a[i] - b[i]; break;
    case T::MULTIPLY: for (int64_t i = 0; i < n; ++i) out[i] = a[i] * b[i]; break;
    case T::DIVIDE: for (int64_t i = 0; i < n; ++i) out[i] = a[i] / b[i]; break;
    case T::LT: for (int64_t i = 0; i < n; ++i) out[i] = a[i] < b[i] ? 1 : 0; break;
    case T::GT: for (int64_t i = 0; i < n; ++i) out[i] = a[i] > b[i] ? 1 : 0; break;
    case T::EQ: for (int64_t i = 0; i < n; ++i) out[i] = a[i] == b[i] ? 1 : 0; break;
    case T::NEQ: for (int64_t i = 0; i < n; ++i) out[i] = a[i] != b[i] ? 1 : 0; break;
    case T::MIN: for (int64_t i = 0; i < n; ++i) out[i] = a[i] < b[i] ? a[i] : b[i]; break;
    case T::MAX: for (int64_t i = 0; i < n; ++i) out[i] = a[i] > b[i] ? a[i] : b[i]; break;
    default: throw std::runtime_error("ERROR: Unrecognized operation.");
  }
  return out;
}

// ------------------------------------------------------------ chunked column-pivoted QR
// The canonical least-squares algorithm for up to TS_MAX_COLS columns: the
// column-pivoted Householder QR of ColPivHouseholderQR.h:482-607 carried out
// on the triangular factor of the data instead of on the n x k data itself.
//   1. Initial norms: ||A_j|| in the chunked order (rrm::chunk_sum), as
//      below.  A column whose norm is Inf is pivoted at Eigen's step 0 (or a
//      NaN-norm column at position 0 is), whose reflector is NaN: the outcome
//      is all-NaN coefficients.  A NaN-norm column never moves (a NaN norm
//      never compares greater, and a swap moves only the column at the step's
//      position), so it is pivoted, with the same outcome, exactly when the
//      factorisation reaches its position before the rank cut; otherwise its
//      coefficient is zero.  The finite columns take part in steps 2-4.
//   2. Per TS_ROWS-row chunk of [finite columns | y]: unpivoted Householder
//      QR (Eigen's makeHouseholder), its triangular factor R_c.  Every sum in
//      a chunk is ts_dot's order.  A reflector (tau, essential part x / den)
//      is applied to a column t as w = t_j + (x . t) / den, t_j -= tau w,
//      t_i -= x_i g with g = (tau w) / den for the rows below: one division
//      per column, and every product-sum and update a fused multiply-add.
//   3. The chunk factors combined in two levels: the chunks fall into groups
//      of TS_GROUP consecutive chunks; within a group, in chunk order, G = the
//      group's first factor, then G = the triangular factor of [G; R_c]
//      (Householder on each column of the stacked pair, sums in row order,
//      the same update); then R = G_0 and the group factors folded into it
//      the same way in group order.  The groups fold independently of each
//      other, so the sequential depth is TS_GROUP + groups instead of the
//      chunk count.
//   4. Eigen's pivoted steps on R: pivot = largest current norm (left to
//      right, strict >), rank cut at biggest^2 < (max initial norm * eps)^2 /
//      n * (n - p), makeHouseholder and its application to the remaining
//      columns and to R's y column; current norms recomputed from R's rows
//      below the step (in row order) instead of downdated.  Then the
//      triangular solve of the leading r x r block, zero for the rest.
// In exact arithmetic every pivot, cut and coefficient is the one of the
// pivoted QR on the data; rounding is this order's.
constexpr int64_t TS_ROWS = 4096;      // rows per chunk
constexpr int TS_MAX_COLS = 9;         // columns (free term included) handled this way
constexpr int64_t TS_GROUP = 16;       // chunks per group of the two-level combination

// The column bucket of a candidate with k columns: columns with y of its
// chunk factors (3, 5 or 10), and the threads of its chunk sums: 256 threads
// of 16 rows for the two small buckets, 1024 of 4 rows for the wide one.
static int ts_bucket(int k) { return k <= 2 ? 3 : k <= 4 ? 5 : 10; }
static int ts_threads(int nb) { return nb <= 5 ? 256 : 1024; }

// Sum of x[i] y[i] over the chunk-local rows lo <= i < m (<= TS_ROWS) with T
// threads: per thread t its rows t + T u (u = 0 .. TS_ROWS/T - 1) in order,
// each accumulated by a fused multiply-add, each group of 64 threads combined
// by the xor butterfly of strides 1, 2, 4, ..., 32 (lane l adds lane l ^ s),
// the T/64 group sums in order.
static double ts_dot(int T, int64_t m, int64_t lo, const double* x, const double* y) {
  double total = 0.0;
  for (int w = 0; w < T / 64; ++w) {
    double a[64];
    for (int l = 0; l < 64; ++l) {
      const int t = 64 * w + l;
      double p = 0.0;
      for (int u = 0; u < (int)(TS_ROWS / T); ++u) {
        const int64_t r = t + (int64_t)T * u;
        if (r >= lo && r < m) p = std::fma(x[r], y[r], p);
      }
      a[l] = p;
    }
    for (int s = 1; s <= 32; s <<= 1) {
      double nx[64];
      for (int l = 0; l < 64; ++l) nx[l] = a[l] + a[l ^ s];
      for (int l = 0; l < 64; ++l) a[l] = nx[l];
    }
    total = (w == 0) ? a[0] : total + a[0];
  }
  return total;
}

// Eigen's makeHouseholder from the pivot c0 and the tail's sum of squares.
static void householder(double c0, double tail_sq, double* beta, double* tau, double* den) {
  if (tail_sq <= 2.2250738585072014e-308) {
    *tau = 0.0; *beta = c0; *den = 1.0;
  } else {
    double b = std::sqrt(c0 * c0 + tail_sq);
    if (c0 >= 0.0) b = -b;
    *beta = b;
    *den = c0 - b;
    *tau = (b - c0) / b;
  }
}

static std::vector<double> ts_qr_solve(const std::vector<std::vector<double>>& A, const std::vector<double>& b,
                                       int64_t n, int* rank_out) {
  const int k = (int)A.size();
  const double eps = 2.220446049250313e-16;
  const int size = (int)std::min<int64_t>(n, k);
  std::vector<double> nd(k);
  for (int j = 0; j < k; ++j) {
    const std::vector<double>& a = A[j];
    nd[j] = std::sqrt(rrm::chunk_sum(0, n, [&](int64_t i) { return a[i] * a[i]; }));
  }
  std::vector<double> coefs(k, 0.0);
  const double inf = std::numeric_limits<double>::infinity();
  std::vector<int> fin;   // finite columns in column order
  for (int j = 0; j < k; ++j)
    if (nd[j] == nd[j] && nd[j] != inf) fin.push_back(j);
  std::vector<int> perm(k);   // position -> column
  std::iota(perm.begin(), perm.end(), 0);
  // a reflector from a non-finite column is NaN; nothing is cut after it and
  // every coefficient of the first `size` positions is NaN
  auto all_nan = [&]() {
    std::vector<double> out(k, 0.0);
    for (int i = 0; i < size; ++i) out[perm[i]] = rrm::qnan();
    *rank_out = size;
    return out;
  };
  {   // step 0's pivot: non-finite when column 0's norm is NaN or any norm is Inf
    int bi = 0;
    double best = nd[0];
    for (int j = 1; j < k; ++j)
      if (nd[j] > best) { best = nd[j]; bi = j; }
    if (!(best == best) || best == inf) {
      std::swap(perm[0], perm[bi]);
      return all_nan();
    }
  }
  const int kf = (int)fin.size();
  const int nc = kf + 1;   // finite columns and y
  const int nthr = ts_threads(ts_bucket(k));
  // 2. chunk factors; 3. their combination.  R: kf rows x nc columns, R[i][c].
  std::vector<double> R((size_t)kf * nc, 0.0), T((size_t)kf * nc);
  // [Rm; Tm] -> Rm, Tm's column j has rows 0..j (Tm is overwritten)
  auto fold = [&](std::vector<double>& Rm, std::vector<double>& Tm) {
    for (int j = 0; j < kf; ++j) {
      double sxx = 0.0;
      for (int i = 0; i <= j; ++i) sxx = std::fma(Tm[(size_t)i * nc + j], Tm[(size_t)i * nc + j], sxx);
      double beta, tau, den;
      householder(Rm[(size_t)j * nc + j], sxx, &beta, &tau, &den);
      if (tau != 0.0) {
        for (int c = j + 1; c < nc; ++c) {
          double sx = 0.0;
          for (int i = 0; i <= j; ++i) sx = std::fma(Tm[(size_t)i * nc + j], Tm[(size_t)i * nc + c], sx);
          const double w = Rm[(size_t)j * nc + c] + sx / den;
          const double g = (tau * w) / den;
          Rm[(size_t)j * nc + c] = Rm[(size_t)j * nc + c] - tau * w;
          for (int i = 0; i <= j; ++i)
            Tm[(size_t)i * nc + c] = std::fma(-Tm[(size_t)i * nc + j], g, Tm[(size_t)i * nc + c]);
        }
      }
      Rm[(size_t)j * nc + j] = beta;
    }
  };
  const int64_t nch = (n + TS_ROWS - 1) / TS_ROWS;
  std::vector<std::vector<double>> G((size_t)((nch + TS_GROUP - 1) / TS_GROUP));
  for (int64_t ch = 0; ch < nch; ++ch) {
    const int64_t r0 = ch * TS_ROWS, m = std::min<int64_t>(TS_ROWS, n - r0);
    std::vector<std::vector<double>> P(nc, std::vector<double>((size_t)m));
    for (int c = 0; c < nc; ++c) {
      const std::vector<double>& src = c < kf ? A[fin[c]] : b;
      for (int64_t i = 0; i < m; ++i) P[c][i] = src[r0 + i];
    }
    const int steps = (int)std::min<int64_t>(kf, m);
    for (int j = 0; j < steps; ++j) {
      const std::vector<double>& x = P[j];
      const double c0 = x[j];
      const double sxx = ts_dot(nthr, m, j + 1, x.data(), x.data());
      std::vector<double> sx(nc, 0.0);
      for (int c = j + 1; c < nc; ++c) sx[c] = ts_dot(nthr, m, j + 1, x.data(), P[c].data());
      double beta, tau, den;
      householder(c0, sxx, &beta, &tau, &den);
      if (tau != 0.0) {
        for (int c = j + 1; c < nc; ++c) {
          std::vector<double>& t = P[c];
          const double w = t[j] + sx[c] / den;
          const double g = (tau * w) / den;
          t[j] = t[j] - tau * w;
          for (int64_t i = j + 1; i < m; ++i) t[i] = std::fma(-x[i], g, t[i]);
        }
      }
      P[j][j] = beta;
    }
    for (int i = 0; i < kf; ++i)
      for (int c = 0; c < nc; ++c) T[(size_t)i * nc + c] = (i < m && c >= i) ? P[c][i] : 0.0;
    std::vector<double>& Gg = G[(size_t)(ch / TS_GROUP)];
    if (ch % TS_GROUP == 0) Gg = T;
    else fold(Gg, T);
  }
  if (!G.empty()) {
    R = G[0];
    for (size_t g = 1; g < G.size(); ++g) fold(R, G[g]);
  }
  // 4. pivoted steps on R
  std::vector<int> fidx(k, -1);   // column -> its column in R
  for (int f = 0; f < kf; ++f) fidx[fin[f]] = f;
  double mx = nd[0];
  for (int j = 1; j < k; ++j) mx = (mx < nd[j]) ? nd[j] : mx;
  const double thr = ((mx * eps) * (mx * eps)) / (double)n;
  std::vector<double> nu(nd);   // current norms by position
  int nonzero = size;
  auto at = [&](int i, int c) -> double& { return R[(size_t)i * nc + c]; };
  for (int p = 0; p < size; ++p) {
    if (fidx[perm[p]] < 0) return all_nan();   // a NaN-norm column reached before the cut
    int bi = p;
    double best = nu[p];
    for (int j = p + 1; j < k; ++j)
      if (nu[j] > best) { best = nu[j]; bi = j; }
    if (best * best < thr * (double)(n - p)) { nonzero = p; break; }
    std::swap(perm[p], perm[bi]);
    std::swap(nu[p], nu[bi]);
    const int f = fidx[perm[p]];
    double sxx = 0.0;
    for (int i = p + 1; i < kf; ++i) sxx = std::fma(at(i, f), at(i, f), sxx);
    double beta, tau, den;
    householder(at(p, f), sxx, &beta, &tau, &den);
    // the remaining finite columns (positions > p) and y
    for (int q = p + 1; q <= k; ++q) {
      const int g = q < k ? fidx[perm[q]] : kf;
      if (g < 0) continue;
      if (tau != 0.0) {
        double sx = 0.0;
        for (int i = p + 1; i < kf; ++i) sx = std::fma(at(i, f), at(i, g), sx);
        const double w = at(p, g) + sx / den;
        const double gg = (tau * w) / den;
        at(p, g) = at(p, g) - tau * w;
        for (int i = p + 1; i < kf; ++i) at(i, g) = std::fma(-at(i, f), gg, at(i, g));
      }
      if (q < k) {
        double s = 0.0;
        for (int i = p + 1; i < kf; ++i) s = std::fma(at(i, g), at(i, g), s);
        nu[q] = std::sqrt(s);
      }
    }
    at(p, f) = beta;
  }
  *rank_out = nonzero;
  if (nonzero == 0) return coefs;
  std::vector<double> c(nonzero);
  for (int i = 0; i < nonzero; ++i) c[i] = at(i, kf);
  for (int i = nonzero - 1; i >= 0; --i) {
    c[i] = c[i] / at(i, fidx[perm[i]]);
    for (int l = 0; l < i; ++l) c[l] = std::fma(-c[i], at(l, fidx[perm[i]]), c[l]);
  }
  for (int i = 0; i < nonzero; ++i) coefs[perm[i]] = c[i];
  for (int j = 0; j < k; ++j)
    if (coefs[j] != coefs[j]) coefs[j] = rrm::qnan();
  return coefs;
}

// ------------------------------------------------------------ column-pivoted QR solve
// A: k columns of n rows (consumed), b: n.  Returns coefficients in column
// order and the number of nonzero pivots.  Up to TS_MAX_COLS columns: the
// chunked form above; more: Eigen's algorithm on the data itself, below.
static std::vector<double> colpiv_qr_data(std::vector<std::vector<double>> A, const std::vector<double>& b,
                                          int64_t n, int* rank_out);
// Set for timing only (orc_set_qr_direct): every solve takes Eigen's order on
// the data, the reference's own algorithm and cost, instead of the canonical
// chunked form the GPU matches bit for bit.
static std::atomic<int> g_qr_direct{0};
static std::vector<double> colpiv_qr_solve(std::vector<std::vector<double>> A, const std::vector<double>& b,
                                           int64_t n, int* rank_out) {
  if ((int)A.size() <= TS_MAX_COLS && !g_qr_direct.load(std::memory_order_relaxed)) return ts_qr_solve(A, b, n, rank_out);
  return colpiv_qr_data(std::move(A), b, n, rank_out);
}

static std::vector<double> colpiv_qr_data(std::vector<std::vector<double>> A, const std::vector<double>& b,
                                          int64_t n, int* rank_out) {
  const int cols = (int)A.size();
  const int64_t rows = n;
  const int size = (int)std::min<int64_t>(rows, cols);
  const double eps = 2.220446049250313e-16;    // NumTraits<double>::epsilon()
  std::vector<double> hc(size), normU(cols), normD(cols);
  std::vector<int> transp(size);
  // every sum over rows in the chunked order (rrm::chunk_sum)
  for (int j = 0; j < cols; ++j) {
    const std::vector<double>& a = A[j];
    normD[j] = std::sqrt(rrm::chunk_sum(0, rows, [&](int64_t i) { return a[i] * a[i]; }));
    normU[j] = normD[j];
  }
  double mx = normU[0];
  for (int j = 1; j < cols; ++j) mx = (mx < normU[j]) ? normU[j] : mx;
  const double threshold_helper = ((mx * eps) * (mx * eps)) / (double)rows;
  const double norm_downdate_threshold = std::sqrt(eps);
  int nonzero = size;
  for (int k = 0; k < size; ++k) {
    int bi = k;
    double best = normU[k];
    for (int j = k + 1; j < cols; ++j)
      if (normU[j] > best) { best = normU[j]; bi = j; }
    const double sq = best * best;
    if (nonzero == size && sq < threshold_helper * (double)(rows - k)) nonzero = k;
    transp[k] = bi;
    if (k != bi) {
      std::swap(A[k], A[bi]);
      std::swap(normU[k], normU[bi]);
      std::swap(normD[k], normD[bi]);
    }
    // makeHouseholderInPlace on A[k][k..rows)
    std::vector<double>& v = A[k];
    const double c0 = v[k];
    const double tailSq = (rows - k == 1) ? 0.0 : rrm::chunk_sum(k + 1, rows, [&](int64_t i) { return v[i] * v[i]; });
    double tau, beta;
    if (tailSq <= 2.2250738585072014e-308) {
      tau = 0.0; beta = c0;
      for (int64_t i = k + 1; i < rows; ++i) v[i] = 0.0;
    } else {
      beta = std::sqrt(c0 * c0 + tailSq);
      if (c0 >= 0.0) beta = -beta;
      const double den = c0 - beta;
      for (int64_t i = k + 1; i < rows; ++i) v[i] = v[i] / den;
      tau = (beta - c0) / beta;
    }
    v[k] = beta;
    hc[k] = tau;
    // applyHouseholderOnTheLeft to columns k+1.. over rows k..rows-1
    if (rows - k == 1) {
      for (int j = k + 1; j < cols; ++j) A[j][k] *= (1.0 - tau);
    } else if (tau != 0.0) {
      for (int j = k + 1; j < cols; ++j) {
        std::vector<double>& a = A[j];
        double t = rrm::chunk_sum(k + 1, rows, [&](int64_t i) { return v[i] * a[i]; });
        t = t + a[k];
        a[k] = a[k] - tau * t;
        for (int64_t i = k + 1; i < rows; ++i) a[i] = a[i] - (tau * v[i]) * t;
      }
    }
    // LAWN-176 norm downdate
    for (int j = k + 1; j < cols; ++j) {
      if (normU[j] != 0.0) {
        double temp = std::fabs(A[j][k]) / normU[j];
        temp = (1.0 + temp) * (1.0 - temp);
        temp = temp < 0.0 ? 0.0 : temp;
        const double ratio = normU[j] / normD[j];
        const double temp2 = temp * (ratio * ratio);
        if (temp2 <= norm_downdate_threshold) {
          const std::vector<double>& a = A[j];
          normD[j] = std::sqrt(rrm::chunk_sum(k + 1, rows, [&](int64_t i) { return a[i] * a[i]; }));
          normU[j] = normD[j];
        } else {
          normU[j] *= std::sqrt(temp);
        }
      }
    }
  }
  std::vector<int> perm(cols);
  std::iota(perm.begin(), perm.end(), 0);
  for (int k = 0; k < size; ++k) std::swap(perm[k], perm[transp[k]]);

  std::vector<double> coefs(cols, 0.0);
  *rank_out = nonzero;
  if (nonzero == 0) return coefs;
  std::vector<double> c = b;
  for (int h = 0; h < nonzero; ++h) {
    const std::vector<double>& v = A[h];
    const double tau = hc[h];
    if (rows - h == 1) {
      c[h] *= (1.0 - tau);
    } else if (tau != 0.0) {
      double t = rrm::chunk_sum(h + 1, rows, [&](int64_t i) { return v[i] * c[i]; });
      t = t + c[h];
      c[h] = c[h] - tau * t;
      for (int64_t i = h + 1; i < rows; ++i) c[i] = c[i] - (tau * v[i]) * t;
    }
  }
  // upper-triangular solve on the leading nonzero x nonzero block, column oriented
  for (int i = nonzero - 1; i >= 0; --i) {
    c[i] = c[i] / A[i][i];
    for (int l = 0; l < i; ++l) c[l] = c[l] - c[i] * A[i][l];
  }
  for (int i = 0; i < nonzero; ++i) coefs[perm[i]] = c[i];
  for (int j = 0; j < cols; ++j)
    if (coefs[j] != coefs[j]) coefs[j] = rrm::qnan();
  return coefs;
}

// ------------------------------------------------------------ fitness
struct Fit { double e0, e1; int size; };
static bool fit_eq(const Fit& a, const Fit& b) { return a.e0 == b.e0 && a.e1 == b.e1 && a.size == b.size; }

static Fit score(const Node& tree, const Data& D) {
  const std::vector<double> yp = evaluate(tree, D);
  const std::vector<double>& y = D.y;
  // the chunked order (rrm::chunk_sum): the GPU scores a candidate's rows in
  // 4096-row chunks on separate workgroups
  const double ssr = rrm::chunk_sum(0, D.n, [&](int64_t i) { return (y[i] - yp[i]) * (y[i] - yp[i]); });
  const double r2 = 1 - ssr / D.sst;
  const double rmse = std::sqrt(ssr / (double)D.n);
  if (r2 != r2 || rmse != rmse) return Fit{1000, 1000, 1000};
  return Fit{1 - r2, rmse, size(tree)};
}

struct Tuned {
  P ols;
  std::vector<double> coefs;
  std::vector<std::string> factors;  // prefix strings, free term excluded
  int rank = 0;
};

// rils_rols_cpp.cpp:445-518.  Mutates `sol` (expand, simplify) like the reference.
static Tuned tune_constants(Node& sol, const Data& D) {
  expand(sol);
  simplify(sol);
  std::vector<Node*> all;
  extract_factors(sol, all);
  std::vector<const Node*> factors;
  for (Node* f : all) {
    if (f->type == T::CONST) continue;
    if (f->ar() == 2 && f->l->type == T::CONST && f->r->type == T::CONST) continue;
    if (f->type == T::MULTIPLY || f->type == T::PLUS || f->type == T::MINUS) {
      if (f->l->type == T::CONST) { factors.push_back(f->r.get()); continue; }
      if (f->r->type == T::CONST) { factors.push_back(f->l.get()); continue; }
    }
    if (f->type == T::DIVIDE && f->r->type == T::CONST) { factors.push_back(f->l.get()); continue; }
    factors.push_back(f);
  }
  Node free_term(T::CONST);
  free_term.c = 1.0;
  factors.push_back(&free_term);

  std::vector<std::vector<double>> A;
  A.reserve(factors.size());
  for (const Node* f : factors) A.push_back(evaluate(*f, D));
  Tuned out;
  out.coefs = colpiv_qr_solve(std::move(A), D.y, D.n, &out.rank);
  for (size_t i = 0; i + 1 < factors.size(); ++i) out.factors.push_back(prefix(*factors[i]));

  P ols;
  for (size_t i = 0; i < factors.size(); ++i) {
    const double coef = out.coefs[i];
    if (value_zero(coef)) continue;
    P nf;
    if (factors[i]->type == T::CONST) nf = mkc(coef * factors[i]->c);
    else if (value_one(coef)) nf = factors[i]->clone();
    else nf = mk2(T::MULTIPLY, mkc(coef), factors[i]->clone());
    ols = ols ? mk2(T::PLUS, std::move(ols), std::move(nf)) : std::move(nf);
  }
  if (!ols) ols = mkc(0.0);
  out.ols = std::move(ols);
  return out;
}

// ------------------------------------------------------------ search
struct Move { int fit_calls; int kind; std::string model; double e0, e1; int size; };

class RilsRols {
 public:
  bool classification = false, verbose = false;
  int max_fit_calls = 100000, max_seconds = 100, random_state = 0, max_feat = 200;
  double complexity_penalty = 0.001, sample_size = 1, max_complexity = 50;

  int main_it = 0, fit_calls = 0, ls_calls = 0, skipped_perts = 0, total_perts = 0;
  std::unordered_set<std::string> checked_perts;
  std::chrono::steady_clock::time_point start_time;
  P final_solution;
  Fit final_fitness{0, 0, 0};
  double best_time = 0, total_time = 0;
  std::vector<P> allowed;
  std::vector<Move> moves;

  void reset() {
    main_it = 0; fit_calls = 0; ls_calls = 0;
    start_time = std::chrono::steady_clock::now();
    checked_perts.clear();
    skipped_perts = 0; total_perts = 0;
    srand(random_state);
  }

  void setup_nodes(const std::vector<int>& rel) {
    const T ops[] = {T::PLUS, T::MINUS, T::MULTIPLY, T::DIVIDE, T::SIN, T::COS, T::LN, T::EXP, T::SQRT, T::SQR};
    for (T t : ops) allowed.push_back(mk(t));
    const double consts[] = {-1., 0., 0.5, 1., 2., PI_, 10.};
    for (double c : consts) allowed.push_back(mkc(c));
    for (int v : rel) allowed.push_back(mkv(v));
    if (classification) {
      const T cops[] = {T::LT, T::GT, T::EQ, T::NEQ, T::MIN, T::MAX};
      for (T t : cops) allowed.push_back(mk(t));
    }
  }

  // generators, rils_rols_cpp.cpp:163-316
  void add_const_finetune(const Node& o, std::vector<P>& c) {
    if (o.type != T::CONST) return;
    if (o.c == 0.0) { c.push_back(mkc(-1.0)); c.push_back(mkc(1.0)); return; }
    const double mult[] = {-1., 0.01, 0.1, 0.2, 0.5, 0.8, 0.9, 0., 1, 1.1, 1.2, 2., PI_, 5., 10., 20., 50., 100.};
    for (double m : mult) c.push_back(mkc(o.c * m));
  }
  void add_change_to_subtree(const Node& o, std::vector<P>& c) {
    if (o.ar() >= 1) all_subtrees(*o.l, c);
    if (o.ar() >= 2) all_subtrees(*o.r, c);
  }
  void add_change_to_var_or_1(const Node& o, std::vector<P>& c) {
    for (const P& n : allowed) {
      if (n->type != T::VAR) continue;
      if (o.type == T::VAR && o.var == n->var) continue;
      c.push_back(n->clone());
    }
    c.push_back(mkc(1.0));
  }
  void add_change_unary_applied(const Node& o, std::vector<P>& c) {
    for (const P& n : allowed) {
      if (n->ar() != 1 || !allowed_left(n->type, o.type)) continue;
      c.push_back(mk1(n->type, o.clone()));
    }
  }
  void add_change_variable_to_unary_applied(const Node& o, std::vector<P>& c) {
    if (o.type == T::VAR) add_change_unary_applied(o, c);
  }
  void add_change_unary_to_another(const Node& o, std::vector<P>& c) {
    if (o.ar() != 1) return;
    for (const P& n : allowed) {
      if (n->ar() != 1 || n->type == o.type) continue;
      if (!allowed_left(n->type, o.l->type)) continue;
      c.push_back(mk1(n->type, o.l->clone()));
    }
  }
  void add_change_binary_applied(const Node& o, std::vector<P>& c) {
    std::vector<P> args;
    all_subtrees(o, args);
    for (const P& n : allowed)
      if (n->type == T::VAR || n->type == T::CONST) args.push_back(n->clone());
    for (const P& nb : allowed) {
      if (nb->ar() != 2) continue;
      for (const P& a : args) {
        if (allowed_left(nb->type, o.type)) c.push_back(mk2(nb->type, o.clone(), a->clone()));
        if (!symmetric(nb->type) && allowed_left(nb->type, a->type)) c.push_back(mk2(nb->type, a->clone(), o.clone()));
      }
    }
  }
  void add_change_variable_constant_to_binary_applied(const Node& o, std::vector<P>& c) {
    if (o.type == T::VAR || o.type == T::CONST) add_change_binary_applied(o, c);
  }
  void add_change_binary_to_another(const Node& o, std::vector<P>& c) {
    if (o.ar() != 2) return;
    for (const P& nb : allowed) {
      if (nb->ar() != 2 || nb->type == o.type) continue;
      if (allowed_left(nb->type, o.l->type)) c.push_back(mk2(nb->type, o.l->clone(), o.r->clone()));
      if (!symmetric(nb->type) && allowed_left(nb->type, o.r->type)) c.push_back(mk2(nb->type, o.r->clone(), o.l->clone()));
    }
  }
  std::vector<P> perturb_candidates(const Node& o) {   // 318-330
    std::vector<P> c;
    add_change_to_subtree(o, c);
    add_change_to_var_or_1(o, c);
    add_change_variable_to_unary_applied(o, c);
    add_change_unary_to_another(o, c);
    add_change_variable_constant_to_binary_applied(o, c);
    add_change_binary_to_another(o, c);
    return c;
  }
  std::vector<P> change_candidates(const Node& o) {    // 332-346
    std::vector<P> c;
    add_const_finetune(o, c);
    add_change_to_subtree(o, c);
    add_change_to_var_or_1(o, c);
    add_change_unary_applied(o, c);
    add_change_unary_to_another(o, c);
    add_change_binary_applied(o, c);
    add_change_binary_to_another(o, c);
    return c;
  }

  // 348-443
  std::vector<P> all_candidates(const Node& passed, bool local_search) {
    P solution = passed.clone();
    std::vector<P> all_cand;
    std::unordered_set<std::string> all_cand_str;
    auto get = [&](const Node& t) { return local_search ? change_candidates(t) : perturb_candidates(t); };
    auto push_unique = [&](P cand) {
      std::string s = to_string(*cand);
      if (all_cand_str.find(s) == all_cand_str.end()) {
        all_cand_str.insert(s);
        all_cand.push_back(std::move(cand));
      }
    };
    std::vector<Node*> subs;
    subs.push_back(solution.get());
    for (size_t i = 0; i < subs.size(); ++i) {
      Node* st = subs[i];
      if (i == 0) {
        std::vector<P> cands = get(*st);
        for (P& c : cands) push_unique(std::move(c));
      }
      if (st->ar() >= 1) {
        std::vector<P> cands = get(*st->l);
        P saved = std::move(st->l);
        for (P& c : cands) { st->l = std::move(c); push_unique(solution->clone()); }
        st->l = std::move(saved);
        subs.push_back(st->l.get());
      }
      if (st->ar() >= 2) {
        std::vector<P> cands = get(*st->r);
        P saved = std::move(st->r);
        for (P& c : cands) { st->r = std::move(c); push_unique(solution->clone()); }
        st->r = std::move(saved);
        subs.push_back(st->r.get());
      }
    }
    for (P& nd : all_cand) {
      int it_max = 5;
      while (it_max > 0) {
        const int sz = size(*nd);
        expand(*nd);
        normalize_factor_constants(*nd, T::NONE, false);
        simplify(*nd);
        if (size(*nd) == sz) break;
        --it_max;
      }
    }
    std::unordered_set<std::string> seen;
    std::vector<P> filtered;
    for (P& nd : all_cand) {
      std::string s = to_string(*nd);
      if (seen.find(s) != seen.end()) continue;
      seen.insert(s);
      filtered.push_back(std::move(nd));
    }
    return filtered;
  }

  bool finished() const {
    const auto el = std::chrono::duration_cast<std::chrono::seconds>(std::chrono::steady_clock::now() - start_time).count();
    return fit_calls >= max_fit_calls || el > max_seconds;
  }
  bool check_skip(const std::string& s) {
    total_perts++;
    if (checked_perts.find(s) != checked_perts.end()) { skipped_perts++; return true; }
    checked_perts.insert(s);
    return false;
  }

  Fit fitness(const Node& t, const Data& D) { fit_calls++; return score(t, D); }
  double fitness_value(const Fit& f) const { return (1 + f.e0) * (1 + f.e1) * (1 + f.size * complexity_penalty); }
  int compare_fitness(const Fit& a, const Fit& b) const {
    if ((a.size > max_complexity || b.size > max_complexity) && a.size != b.size) return a.size - b.size;
    const double x = fitness_value(a), y = fitness_value(b);
    if (x < y) return -1;
    if (x > y) return 1;
    return 0;
  }
  static bool dominates(const Fit& p, const Fit& f) { return p.e0 <= f.e0 && p.e1 <= f.e1 && p.size <= f.size; }
  static bool is_dominated(const std::vector<Fit>& par, const Fit& f) {
    for (const Fit& p : par) if (dominates(p, f)) return true;
    return false;
  }
  static void add_to_pareto(std::vector<Fit>& par, const Fit& f) {
    if (is_dominated(par, f)) return;
    for (int i = (int)par.size() - 1; i >= 0; --i)
      if (dominates(f, par[i])) par.erase(par.begin() + i);
    par.push_back(f);
  }

  void log_move(int kind, const Node& t, const Fit& f) { moves.push_back(Move{fit_calls, kind, to_string(t), f.e0, f.e1, f.size}); }

  P local_search(const Node& passed, const Data& D) {   // 601-643
    std::vector<Fit> pareto;
    ls_calls++;
    bool improved = true;
    P curr = passed.clone();
    curr = tune_constants(*curr, D).ols;
    Fit curr_fit = fitness(*curr, D);
    while (improved && !finished()) {
      improved = false;
      std::vector<P> perts = all_candidates(*curr, true);
      for (size_t j = 0; j < perts.size(); ++j) {
        if (finished()) break;
        P lp = perts[j]->clone();
        P tuned = tune_constants(*lp, D).ols;
        Fit tf = fitness(*tuned, D);
        if (!is_dominated(pareto, tf) && compare_fitness(tf, curr_fit) < 0) {
          improved = true;
          int it_max = 5;
          while (it_max > 0) {
            const int sz = size(*tuned);
            expand(*tuned);
            simplify(*tuned);
            if (sz == size(*tuned)) break;
            --it_max;
          }
          tf = fitness(*tuned, D);
          curr = std::move(tuned);
          curr_fit = tf;
          add_to_pareto(pareto, tf);
          log_move(0, *curr, curr_fit);
        }
      }
    }
    return curr;
  }

  std::vector<int> relevant_features(const Data& D) {   // 753-770
    std::vector<int> rel;
    if (D.d <= max_feat) { rel.resize(D.d); std::iota(rel.begin(), rel.end(), 0); return rel; }
    std::vector<std::tuple<double, int>> by;
    for (int i = 0; i < D.d; ++i) {
      // R2(X[i], y): mean and total sum of squares are taken from X[i]
      const std::vector<double>& x = D.X[i];
      const std::vector<double>& y = D.y;
      const double ssr = canon_sum(0, D.n, [&](int64_t r) { return (x[r] - y[r]) * (x[r] - y[r]); });
      Data xi; xi.n = D.n; xi.y = x; finish_data(xi);
      by.push_back(std::tuple<double, int>{1 - ssr / xi.sst, i});
    }
    std::sort(by.begin(), by.end(), std::greater<>());
    for (int i = 0; i < max_feat; ++i) rel.push_back(std::get<1>(by[i]));
    return rel;
  }

  void fit_inner(const std::vector<std::vector<double>>& X_all, const std::vector<double>& y_all) {  // 772-859
    reset();
    const int64_t N = (int64_t)y_all.size();
    const int sample_cnt = int(sample_size * (double)N);
    std::vector<int> selected(N);
    std::iota(selected.begin(), selected.end(), 0);
    std::shuffle(selected.begin(), selected.end(), std::default_random_engine(random_state));
    Data D;
    D.n = sample_cnt;
    D.d = (int)X_all.size();
    D.X.assign(X_all.size(), std::vector<double>(sample_cnt));
    D.y.resize(sample_cnt);
    for (int ix = 0; ix < sample_cnt; ++ix) {
      const int i = selected[ix];
      for (size_t j = 0; j < X_all.size(); ++j) D.X[j][ix] = X_all[j][i];
      D.y[ix] = y_all[i];
    }
    finish_data(D);
    setup_nodes(relevant_features(D));
    final_solution = mkc(0.0);
    final_fitness = fitness(*final_solution, D);
    bool improved = true;
    while (!finished()) {
      main_it += 1;
      P start = final_solution->clone();
      if (!improved) {
        std::vector<P> all_perts = all_candidates(*final_solution, false);
        const size_t i1 = (size_t)rand() % all_perts.size();
        std::vector<P> all2 = all_candidates(*all_perts[i1], false);
        const size_t i2 = (size_t)rand() % all2.size();
        start = all2[i2]->clone();
      }
      improved = false;
      std::vector<P> all_perts = all_candidates(*start, false);
      std::vector<std::tuple<double, P>> by_perts;
      for (size_t i = 0; i < all_perts.size(); ++i) {
        if (finished()) break;
        const std::string ps = to_string(*all_perts[i]);
        if (check_skip(ps)) continue;
        P pt = all_perts[i]->clone();
        const Fit f = fitness(*pt, D);
        by_perts.emplace_back(f.e0, std::move(pt));
      }
      std::sort(by_perts.begin(), by_perts.end(),
                [](const std::tuple<double, P>& a, const std::tuple<double, P>& b) { return std::get<0>(a) < std::get<0>(b); });
      for (size_t i = 0; i < by_perts.size(); ++i) {
        if (finished()) break;
        const Node& lp = *std::get<1>(by_perts[i]);
        checked_perts.insert(to_string(lp));
        P res = local_search(lp, D);
        const Fit f = fitness(*res, D);
        if (compare_fitness(f, final_fitness) < 0) {
          improved = true;
          final_solution = res->clone();
          final_fitness = f;
          log_move(1, *final_solution, final_fitness);
          best_time = std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - start_time).count() / 1000.0;
        }
      }
    }
    total_time = std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - start_time).count() / 1000.0;
  }
};

// ------------------------------------------------------------ C API
static thread_local std::string g_err;
static std::vector<Data*> g_data;

static char* dup_str(const std::string& s) {
  char* p = (char*)malloc(s.size() + 1);
  memcpy(p, s.data(), s.size() + 1);
  return p;
}

}  // namespace orc

using namespace orc;

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }
void orc_free(void* p) { free(p); }

// op: 0 sin, 1 cos, 2 exp, 3 log, 4 sqrt
int orc_math(int op, const double* in, double* out, int64_t n) {
  for (int64_t i = 0; i < n; ++i) {
    const double x = in[i];
    switch (op) {
      case 0: out[i] = rrm::sin(x); break;
      case 1: out[i] = rrm::cos(x); break;
      case 2: out[i] = rrm::exp(x); break;
      case 3: out[i] = rrm::log(x); break;
      case 4: out[i] = rrm::sqrt(x); break;
      default: return -1;
    }
  }
  return 0;
}

double orc_chunk_sum(const double* v, int64_t lo, int64_t hi) {
  return rrm::chunk_sum(lo, hi, [&](int64_t i) { return v[i]; });
}

double orc_canon_sum(const double* v, int64_t lo, int64_t hi) {
  return canon_sum(lo, hi, [&](int64_t i) { return v[i]; });
}

// X feature-major (d columns of n)
int orc_dataset(const double* X_fm, const double* y, int64_t n, int d) {
  Data* D = new Data;
  D->n = n; D->d = d;
  D->X.assign(d, std::vector<double>(n));
  for (int j = 0; j < d; ++j) for (int64_t i = 0; i < n; ++i) D->X[j][i] = X_fm[(int64_t)j * n + i];
  D->y.assign(y, y + n);
  finish_data(*D);
  g_data.push_back(D);
  return (int)g_data.size() - 1;
}
void orc_dataset_free(int h) { if (h >= 0 && h < (int)g_data.size()) { delete g_data[h]; g_data[h] = nullptr; } }
int orc_dataset_stats(int h, double* ybar, double* sst) { *ybar = g_data[h]->ybar; *sst = g_data[h]->sst; return 0; }

int orc_eval(int h, const char* pre, double* out) {
  try {
    P t = parse(pre);
    std::vector<double> v = evaluate(*t, *g_data[h]);
    memcpy(out, v.data(), v.size() * sizeof(double));
    return 0;
  } catch (std::exception& e) { g_err = e.what(); return -1; }
}

// fitness() of a tree as given (no OLS): out = {1-R2, RMSE, size}
int orc_score(int h, const char* pre, double* out3) {
  try {
    P t = parse(pre);
    Fit f = score(*t, *g_data[h]);
    out3[0] = f.e0; out3[1] = f.e1; out3[2] = f.size;
    return 0;
  } catch (std::exception& e) { g_err = e.what(); return -1; }
}

// tune_constants + fitness: returns the OLS tree (prefix), coefficients
// (factor order, free term last), rank and fitness.  *factors_out lists the
// factor subtrees (prefix, '\n'-separated).
int orc_tune(int h, const char* pre, char** ols_out, char** factors_out, double* coefs, int cap, int* k_out,
             int* rank_out, double* fit3) {
  try {
    P t = parse(pre);
    Tuned tu = tune_constants(*t, *g_data[h]);
    const int k = (int)tu.coefs.size();
    *k_out = k;
    *rank_out = tu.rank;
    for (int i = 0; i < k && i < cap; ++i) coefs[i] = tu.coefs[i];
    Fit f = score(*tu.ols, *g_data[h]);
    fit3[0] = f.e0; fit3[1] = f.e1; fit3[2] = f.size;
    *ols_out = dup_str(prefix(*tu.ols));
    std::string fs;
    for (size_t i = 0; i < tu.factors.size(); ++i) { if (i) fs += '\n'; fs += tu.factors[i]; }
    *factors_out = dup_str(fs);
    return 0;
  } catch (std::exception& e) { g_err = e.what(); return -1; }
}

// A column-major n x k
void orc_set_qr_direct(int on) { g_qr_direct.store(on ? 1 : 0); }

// direct: 1 to take Eigen's order on the data itself whatever k is (the
// chunked form's cross-check), 0 for the canonical choice by k
int orc_qr_solve(const double* A, const double* b, int64_t n, int k, double* coefs, int* rank, int direct) {
  std::vector<std::vector<double>> cols(k, std::vector<double>(n));
  for (int j = 0; j < k; ++j) for (int64_t i = 0; i < n; ++i) cols[j][i] = A[(int64_t)j * n + i];
  std::vector<double> bb(b, b + n);
  std::vector<double> c = direct ? colpiv_qr_data(std::move(cols), bb, n, rank)
                                 : colpiv_qr_solve(std::move(cols), bb, n, rank);
  for (int j = 0; j < k; ++j) coefs[j] = c[j];
  return 0;
}

// op: 0 to_string (reference format), 1 expand, 2 simplify,
// 3 normalize_factor_constants(NONE,false), 4 size, 5 candidate rewrite
// rounds (all_candidates 413-425), 6 acceptance rewrite rounds (621-630),
// 7 factors (extract + filter of tune_constants, '\n'-separated prefix)
char* orc_rewrite(const char* pre, int op) {
  try {
    P t = parse(pre);
    switch (op) {
      case 0: return dup_str(to_string(*t));
      case 1: expand(*t); break;
      case 2: simplify(*t); break;
      case 3: normalize_factor_constants(*t, T::NONE, false); break;
      case 4: return dup_str(std::to_string(size(*t)));
      case 5: {
        int it_max = 5;
        while (it_max > 0) {
          const int sz = size(*t);
          expand(*t); normalize_factor_constants(*t, T::NONE, false); simplify(*t);
          if (size(*t) == sz) break;
          --it_max;
        }
        break;
      }
      case 6: {
        int it_max = 5;
        while (it_max > 0) {
          const int sz = size(*t);
          expand(*t); simplify(*t);
          if (size(*t) == sz) break;
          --it_max;
        }
        break;
      }
      default: g_err = "bad op"; return nullptr;
    }
    return dup_str(prefix(*t));
  } catch (std::exception& e) { g_err = e.what(); return nullptr; }
}

// Candidates of all_candidates(tree, local_search) for a problem with
// `nvars` features; '\n'-separated prefix strings.
char* orc_all_candidates(const char* pre, int nvars, int classification, int local_search) {
  try {
    RilsRols rr;
    rr.classification = classification != 0;
    std::vector<int> rel(nvars);
    std::iota(rel.begin(), rel.end(), 0);
    rr.setup_nodes(rel);
    P t = parse(pre);
    std::vector<P> c = rr.all_candidates(*t, local_search != 0);
    std::string s;
    for (size_t i = 0; i < c.size(); ++i) { if (i) s += '\n'; s += prefix(*c[i]); }
    return dup_str(s);
  } catch (std::exception& e) { g_err = e.what(); return nullptr; }
}

// Full fit (rils_rols::fit + fit_inner) on row-major X.  Returns a text
// report: first line "model\t<string>", then "fit_calls\t<int>", then one
// line per accepted move "move\t<kind>\t<fit_calls>\t<e0 %a>\t<e1 %a>\t<size>\t<model>".
char* orc_fit(const double* X_rm, const double* y, int n, int d, int classification, int max_fit_calls,
              int max_seconds, double complexity_penalty, int max_complexity, double sample_size, int random_state) {
  try {
    RilsRols rr;
    rr.classification = classification != 0;
    rr.max_fit_calls = max_fit_calls;
    rr.max_seconds = max_seconds;
    rr.complexity_penalty = complexity_penalty;
    rr.max_complexity = max_complexity;
    rr.sample_size = sample_size;
    rr.random_state = random_state;
    rr.reset();
    std::vector<std::vector<double>> Xe(d, std::vector<double>(n));
    for (int i = 0; i < n; ++i) for (int j = 0; j < d; ++j) Xe[j][i] = X_rm[(int64_t)i * d + j];
    std::vector<double> ye(y, y + n);
    rr.fit_inner(Xe, ye);
    std::string s = "model\t" + to_string(*rr.final_solution) + "\n";
    s += "prefix\t" + prefix(*rr.final_solution) + "\n";
    s += "fit_calls\t" + std::to_string(rr.fit_calls) + "\n";
    char b[256];
    snprintf(b, sizeof b, "final\t%a\t%a\t%d\n", rr.final_fitness.e0, rr.final_fitness.e1, rr.final_fitness.size);
    s += b;
    for (const Move& m : rr.moves) {
      snprintf(b, sizeof b, "move\t%d\t%d\t%a\t%a\t%d\t", m.kind, m.fit_calls, m.e0, m.e1, m.size);
      s += b;
      s += m.model;
      s += "\n";
    }
    return dup_str(s);
  } catch (std::exception& e) { g_err = e.what(); return nullptr; }
}

// One local_search (rils_rols_cpp.cpp:601-643) from the tree `pre` on
// row-major X (every feature allowed), with a fit-call budget.  Report as
// orc_fit's: model and prefix of the returned solution, fit_calls, the
// fitness of the returned solution (one more fitness call, not counted),
// and one "move" line per local-search acceptance.
char* orc_local_search(const double* X_rm, const double* y, int n, int d, const char* pre, int max_fit_calls,
                       double complexity_penalty, int max_complexity) {
  try {
    RilsRols rr;
    rr.max_fit_calls = max_fit_calls;
    rr.max_seconds = INT32_MAX;
    rr.complexity_penalty = complexity_penalty;
    rr.max_complexity = max_complexity;
    rr.reset();
    Data D;
    D.n = n;
    D.d = d;
    D.X.assign(d, std::vector<double>(n));
    for (int i = 0; i < n; ++i) for (int j = 0; j < d; ++j) D.X[j][i] = X_rm[(int64_t)i * d + j];
    D.y.assign(y, y + n);
    finish_data(D);
    std::vector<int> rel(d);
    std::iota(rel.begin(), rel.end(), 0);
    rr.setup_nodes(rel);
    P t = parse(pre);
    P res = rr.local_search(*t, D);
    const Fit f = score(*res, D);
    std::string s = "model\t" + to_string(*res) + "\n";
    s += "prefix\t" + prefix(*res) + "\n";
    s += "fit_calls\t" + std::to_string(rr.fit_calls) + "\n";
    char b[256];
    snprintf(b, sizeof b, "final\t%a\t%a\t%d\n", f.e0, f.e1, f.size);
    s += b;
    for (const Move& m : rr.moves) {
      snprintf(b, sizeof b, "move\t%d\t%d\t%a\t%a\t%d\t", m.kind, m.fit_calls, m.e0, m.e1, m.size);
      s += b;
      s += m.model;
      s += "\n";
    }
    return dup_str(s);
  } catch (std::exception& e) { g_err = e.what(); return nullptr; }
}

// relevant_features (rils_rols_cpp.cpp:753-770) of row-major X and y: the
// kept feature indices in order into out (at most max(d, 200) of them);
// returns their count.
int orc_relevant_features(const double* X_rm, const double* y, int n, int d, int* out) {
  RilsRols rr;
  Data D;
  D.n = n;
  D.d = d;
  D.X.assign(d, std::vector<double>(n));
  for (int i = 0; i < n; ++i) for (int j = 0; j < d; ++j) D.X[j][i] = X_rm[(int64_t)i * d + j];
  D.y.assign(y, y + n);
  finish_data(D);
  std::vector<int> rel = rr.relevant_features(D);
  for (size_t i = 0; i < rel.size(); ++i) out[i] = rel[i];
  return (int)rel.size();
}

// Timing helper for the CPU baseline: tune_constants + fitness for each
// prefix string in `pres` ('\n'-separated) with `threads` std::threads over
// candidates.  Returns elapsed seconds; fitness written to fit3 (3 per cand).
double orc_batch_tune(int h, const char* pres, int threads, double* fit3) {
  std::vector<std::string> items;
  {
    const char* p = pres;
    while (*p) {
      const char* q = strchr(p, '\n');
      if (!q) { items.emplace_back(p); break; }
      items.emplace_back(p, q - p);
      p = q + 1;
    }
  }
  const Data& D = *g_data[h];
  auto t0 = std::chrono::steady_clock::now();
  auto work = [&](size_t lo, size_t hi) {
    for (size_t i = lo; i < hi; ++i) {
      P t = parse(items[i]);
      Tuned tu = tune_constants(*t, D);
      Fit f = score(*tu.ols, D);
      fit3[3 * i] = f.e0; fit3[3 * i + 1] = f.e1; fit3[3 * i + 2] = f.size;
    }
  };
  if (threads <= 1) work(0, items.size());
  else {
    std::vector<std::thread> th;
    std::atomic<size_t> next{0};
    for (int t = 0; t < threads; ++t)
      th.emplace_back([&]() { size_t i; while ((i = next.fetch_add(1)) < items.size()) work(i, i + 1); });
    for (auto& x : th) x.join();
  }
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// orc_tune over many prefixes ('\n'-separated) on `threads` std::threads:
// per candidate i its coefficient count k_out[i], coefficients at
// coefs[cap*i ...], rank rank_out[i], fitness fit3[3i..3i+2]; *factors_out
// holds the factor programs of every candidate, '\n'-separated within a
// candidate and candidates separated by '\t'.  Returns 0, or -1 (message in
// orc_last_error) when a candidate has more than `cap` coefficients.
int orc_batch_tune_full(int h, const char* pres, int threads, double* coefs, int cap, int* k_out, int* rank_out,
                        double* fit3, char** factors_out) {
  std::vector<std::string> items;
  for (const char* p = pres; *p;) {
    const char* q = strchr(p, '\n');
    if (!q) { items.emplace_back(p); break; }
    items.emplace_back(p, q - p);
    p = q + 1;
  }
  const Data& D = *g_data[h];
  std::vector<std::string> facs(items.size());
  std::atomic<int> bad{-1};
  auto one = [&](size_t i) {
    P t = parse(items[i]);
    Tuned tu = tune_constants(*t, D);
    const int k = (int)tu.coefs.size();
    k_out[i] = k;
    rank_out[i] = tu.rank;
    if (k > cap) { bad = (int)i; return; }
    for (int j = 0; j < k; ++j) coefs[(size_t)cap * i + j] = tu.coefs[j];
    Fit f = score(*tu.ols, D);
    fit3[3 * i] = f.e0; fit3[3 * i + 1] = f.e1; fit3[3 * i + 2] = f.size;
    for (size_t j = 0; j < tu.factors.size(); ++j) { if (j) facs[i] += '\n'; facs[i] += tu.factors[j]; }
  };
  std::atomic<size_t> next{0};
  auto worker = [&]() { size_t i; while ((i = next.fetch_add(1)) < items.size()) one(i); };
  std::vector<std::thread> th;
  for (int t = 1; t < std::max(threads, 1); ++t) th.emplace_back(worker);
  worker();
  for (auto& x : th) x.join();
  if (bad >= 0) { g_err = "candidate " + std::to_string(bad.load()) + " has more coefficients than cap"; return -1; }
  std::string s;
  for (size_t i = 0; i < facs.size(); ++i) { if (i) s += '\t'; s += facs[i]; }
  *factors_out = dup_str(s);
  return 0;
}

// Acceptance primitives of the search (rils_rols_cpp.cpp:543-599) on two
// fitness tuples a = (e0, e1, size), b: out[0] = fitness_value(a),
// out[1] = compare_fitness(a, b), out[2] = dominates(a, b).
int orc_accept(const double* a, const double* b, double complexity_penalty, int max_complexity, double* out) {
  orc::RilsRols rr;
  rr.complexity_penalty = complexity_penalty;
  rr.max_complexity = max_complexity;
  const orc::Fit fa{a[0], a[1], (int)a[2]}, fb{b[0], b[1], (int)b[2]};
  out[0] = rr.fitness_value(fa);
  out[1] = (double)rr.compare_fitness(fa, fb);
  out[2] = orc::RilsRols::dominates(fa, fb) ? 1.0 : 0.0;
  return 0;
}

}  // extern "C"
