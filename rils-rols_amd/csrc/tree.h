// tree.h -- host-side expression engine of the product path.
//
// Trees live in two forms:
//   Flat  -- a prefix-order token array (the form candidates are stored,
//            spliced, compared and compiled in);
//   Pool  -- an index-linked node pool in which the in-place rewrites of the
//            reference (expand, simplify, normalize_factor_constants) run.
// Semantics follow rils_rols_cpp/node.h and node.cpp of the reference,
// including its printing and rewrite quirks; see the .cpp for citations.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace rrh {

enum Ty : uint8_t {
  NONE, CONST, VAR, PLUS, MINUS, MULTIPLY, DIVIDE, SIN, COS, LN, EXP, SQRT, SQR, POW,
  LT, GT, EQ, NEQ, MIN, MAX
};

inline int arity(uint8_t t) {
  if (t == CONST || t == VAR) return 0;
  if (t >= SIN && t <= SQR) return 1;
  return 2;
}
inline bool symmetric(uint8_t t) { return !(t == MINUS || t == DIVIDE || t == POW); }

struct Tok {
  uint8_t t;
  int32_t var;
  double c;
};
using Flat = std::vector<Tok>;

inline Tok tok_const(double c) { return Tok{CONST, -1, c}; }
inline Tok tok_var(int v) { return Tok{VAR, v, 0.0}; }
inline Tok tok_op(uint8_t t) { return Tok{t, -1, 0.0}; }

// Sizes of every subtree of a flat tree (size[i] = extent of subtree at i).
void subtree_sizes(const Flat& f, std::vector<int32_t>& size);

// Reference printing (node.h:255-301) of the whole tree / a subtree range.
std::string to_string(const Flat& f);
void append_string(const Flat& f, int pos, std::string& out, std::vector<int32_t>* begin_end = nullptr);
void append_const(double c, std::string& out);

// Exchange format: space-separated prefix tokens, constants as "c<%a>".
std::string to_prefix(const Flat& f);
Flat parse_prefix(const std::string& s);

// ---- node pool and in-place rewrites
struct PNode {
  uint8_t t;
  int32_t var;
  double c;
  int32_t l, r;
};

class alignas(64) Pool {   // a cache line of its own: each worker thread owns one
 public:
  std::vector<PNode> n;
  void clear() { n.clear(); }
  int32_t load(const Flat& f, int pos = 0, int* next = nullptr);   // returns root
  int32_t clone(int32_t i);
  void store(int32_t root, Flat& out) const;
  int size(int32_t i) const;
  void expand(int32_t i);
  void simplify(int32_t i);
  void normalize_factor_constants(int32_t i, bool inside);
  void extract_factors(int32_t i, std::vector<int32_t>& out) const;
  int32_t mk(uint8_t t, int32_t l, int32_t r);
  int32_t mkc(double c);
 private:
  void update_with(int32_t self, int32_t src);
  int32_t binomial_mult(int32_t left, int32_t right);
};

// Rewrite rounds: expand -> [normalize_factor_constants] -> simplify, up to
// five times while the size changes (rils_rols_cpp.cpp:413-425 with
// normalize, 621-630 without).
void rewrite_rounds(Flat& f, bool normalize, Pool& pool);

// tune_constants' factor split (rils_rols_cpp.cpp:448-473): expand and
// simplify the candidate once, extract top-level +/- terms and filter them.
// Returns the factor subtrees (free term not included).
void ols_factors(const Flat& cand, std::vector<Flat>& factors, Pool& pool);

// The OLS tree of rils_rols_cpp.cpp:488-517 from factors and coefficients
// (coefs has factors.size()+1 entries, the free term last).
void ols_tree(const std::vector<Flat>& factors, const double* coefs, Flat& out);
int ols_tree_size(const std::vector<int32_t>& factor_sizes, const double* coefs);

// ---- bytecode (include/rr_fitness.h) with Sethi-Ullman ordering
struct Program {
  std::vector<uint32_t> code;
  int depth = 0;
};
// Appends the program of subtree `f` to `code`, its constants to `consts`.
// Returns the operand-stack depth it needs.
int compile(const Flat& f, std::vector<uint32_t>& code, std::vector<double>& consts);

}  // namespace rrh
