// search.cpp -- RILS-ROLS search driver with batched GPU fitness (see search.h).
#include "search.h"

#define RR_HD
#include "rr_math.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <iostream>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string_view>
#include <utility>

namespace rrh {

// ------------------------------------------------------------------ CRand
void CRand::seed(unsigned s) {
  memset(state_, 0, sizeof state_);
  memset(&data_, 0, sizeof data_);
  initstate_r(s, state_, sizeof state_, &data_);
  srandom_r(s, &data_);
}
int CRand::next() {
  int32_t r = 0;
  random_r(&data_, &r);
  return (int)r;
}

// ------------------------------------------------------------------ GPU session
static void rr_check(int rc, const char* what) {
  if (rc != RR_OK) throw std::runtime_error(std::string(what) + ": " + rr_last_error());
}

class DirectBackend : public FitnessBackend {
 public:
  rr_ctx* ctx = nullptr;
  DirectBackend(int device, int64_t slab_mb, const std::vector<std::pair<std::string, int64_t>>& options) {
    rr_check(rr_create(device, &ctx), "rr_create (the MI355X fitness library needs a HIP device)");
    if (slab_mb > 0) rr_check(rr_set_option(ctx, "slab_budget_mb", slab_mb), "rr_set_option");
    for (const auto& o : options) rr_check(rr_set_option(ctx, o.first.c_str(), o.second), "rr_set_option");
  }
  ~DirectBackend() override { if (ctx) rr_destroy(ctx); }
  int32_t upload(const double* X_fm, const double* y, int64_t n, int32_t d) override {
    int32_t h = -1;
    rr_check(rr_dataset_upload(ctx, X_fm, y, n, d, &h), "rr_dataset_upload");
    return h;
  }
  void release(int32_t h) override { rr_dataset_free(ctx, h); }
  void fitness(int32_t h, const rr_batch& b, int32_t mode, int32_t precision, double* coefs, rr_fit* fit,
               double* kernel_ms) override {
    rr_check(rr_fitness_batch(ctx, h, &b, mode, precision, coefs, fit), "rr_fitness_batch");
    rr_last_kernel_ms(ctx, kernel_ms);
  }
  rr_ctx* context() override { return ctx; }
  void autotune(int32_t h, const rr_batch& b, int32_t mode, int32_t precision) override {
    double ms = 0;
    rr_check(rr_autotune(ctx, h, &b, mode, precision, 2, &ms), "rr_autotune");
  }
};

class GpuSession {
 public:
  FitnessBackend* be = nullptr;
  bool own = false;
  int32_t ds = -1;
  int64_t n = 0;
  int d = 0;

  GpuSession(FitnessBackend* backend, int device, int64_t slab_mb,
             const std::vector<std::pair<std::string, int64_t>>& options) {
    if (backend) {
      be = backend;
    } else {
      be = new DirectBackend(device, slab_mb, options);
      own = true;
    }
  }
  ~GpuSession() {
    if (ds >= 0) be->release(ds);
    if (own) delete be;
  }
  void upload(const std::vector<std::vector<double>>& X, const std::vector<double>& y) {
    n = (int64_t)y.size();
    d = (int)X.size();
    std::vector<double> fm((size_t)n * std::max(d, 1));
    for (int j = 0; j < d; ++j) std::copy(X[j].begin(), X[j].end(), fm.begin() + (size_t)j * n);
    if (ds >= 0) be->release(ds);
    ds = -1;
    ds = be->upload(fm.data(), y.data(), n, d);
  }

  struct Builder {
    std::vector<int32_t> cand_start{0}, prog_start{0};
    std::vector<uint32_t> code;
    std::vector<double> consts;
    void add(const Flat& f) {
      compile(f, code, consts);
      prog_start.push_back((int32_t)code.size());
    }
    void close_candidate() { cand_start.push_back((int32_t)prog_start.size() - 1); }
    rr_batch view() const {
      rr_batch b;
      b.n_cands = (int32_t)cand_start.size() - 1;
      b.cand_start = cand_start.data();
      b.prog_start = prog_start.data();
      b.code = code.data();
      b.consts = consts.data();
      b.n_consts = (int32_t)consts.size();
      return b;
    }
  };

  void raw(const std::vector<const Flat*>& trees, std::vector<rr_fit>& out, Stats& st) {
    Builder B;
    for (const Flat* t : trees) { B.add(*t); B.close_candidate(); }
    out.resize(trees.size());
    if (trees.empty()) return;
    const rr_batch b = B.view();
    const auto t0 = std::chrono::steady_clock::now();
    double ms = 0;
    be->fitness(ds, b, RR_MODE_RAW, RR_FP64, nullptr, out.data(), &ms);
    st.gpu_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    st.kernel_ms += ms;
    st.raw_batches++;
    st.raw_cands += (int64_t)trees.size();
  }

  // factors[c] are the OLS factors of candidate c; coefficients of c start at coefs[off[c]]
  void ols(const std::vector<std::vector<Flat>>& factors, size_t count, int precision, std::vector<double>& coefs,
           std::vector<int64_t>& off, std::vector<rr_fit>& out, Stats& st) {
    ols_subset(factors, nullptr, count, precision, coefs, off, out, st);
  }
  // the same for the candidates idx[0..count) (all of 0..count when idx is null);
  // results are indexed by position in idx
  bool tune_next = false;   // run the autotuner on the next batch before scoring it
  void ols_subset(const std::vector<std::vector<Flat>>& factors, const int32_t* idx, size_t count, int precision,
                  std::vector<double>& coefs, std::vector<int64_t>& off, std::vector<rr_fit>& out, Stats& st) {
    Builder B;
    off.resize(count);
    int64_t total = 0;
    for (size_t q = 0; q < count; ++q) {
      const std::vector<Flat>& fc = factors[idx ? (size_t)idx[q] : q];
      for (const Flat& f : fc) B.add(f);
      B.close_candidate();
      off[q] = total;
      total += (int64_t)fc.size() + 1;
    }
    coefs.resize(std::max<int64_t>(total, 1));
    out.resize(count);
    if (count == 0) return;
    const rr_batch b = B.view();
    if (tune_next) {
      tune_next = false;
      be->autotune(ds, b, RR_MODE_OLS, precision);
    }
    const auto t0 = std::chrono::steady_clock::now();
    double ms = 0;
    be->fitness(ds, b, RR_MODE_OLS, precision, coefs.data(), out.data(), &ms);
    st.gpu_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    st.kernel_ms += ms;
  }
};

// ------------------------------------------------------------------ basics
static double secs_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

RilsRols::RilsRols(bool classification, int max_fit_calls, int max_seconds, double complexity_penalty,
                   int max_complexity, double sample_size, bool verbose, int random_state)
    : max_fit_calls_(max_fit_calls), max_seconds_(max_seconds), random_state_(random_state),
      complexity_penalty_(complexity_penalty), sample_size_(sample_size), max_complexity_(max_complexity),
      classification_(classification), verbose_(verbose) {
  reset();
}

RilsRols::~RilsRols() {
  delete gpu_;
  delete workers_;
}

void RilsRols::set_host_threads(int t) {
  delete workers_;
  workers_ = nullptr;
  host_threads_ = t < 0 ? 0 : t;
}

Workers& RilsRols::workers() {
  if (!workers_) {
    int t = host_threads_;
    if (t == 0) {
      const unsigned hw = std::thread::hardware_concurrency();
      t = (int)std::min<unsigned>(hw ? hw : 1, 8);
    }
    workers_ = new Workers(t);
    pools_.assign((size_t)workers_->size(), Pool());
    gen_.assign((size_t)workers_->size(), GenScratch());
  }
  return *workers_;
}

void RilsRols::reset() {   // rils_rols_cpp.cpp:106-117
  main_it_ = 0;
  last_improved_it_ = 0;
  fit_calls_ = 0;
  ls_calls_ = 0;
  start_time_ = std::chrono::steady_clock::now();
  checked_perts_.clear();
  skipped_perts_ = 0;
  total_perts_ = 0;
  rng_.seed((unsigned)random_state_);
}

bool RilsRols::finished() const {   // rils_rols_cpp.cpp:659-661
  const auto el = std::chrono::duration_cast<std::chrono::seconds>(std::chrono::steady_clock::now() - start_time_).count();
  return fit_calls_ >= max_fit_calls_ || el > max_seconds_;
}

bool RilsRols::check_skip(const std::string& s) {   // 663-672
  total_perts_++;
  if (checked_perts_.find(s) != checked_perts_.end()) { skipped_perts_++; return true; }
  checked_perts_.insert(s);
  return false;
}

double RilsRols::fitness_value(const Fit& f) const {   // 543-545
  return (1 + f.e0) * (1 + f.e1) * (1 + f.size * complexity_penalty_);
}

int RilsRols::compare_fitness(const Fit& a, const Fit& b) const {   // 547-561
  if ((a.size > max_complexity_ || b.size > max_complexity_) && a.size != b.size) return a.size - b.size;
  const double x = fitness_value(a), y = fitness_value(b);
  if (x < y) return -1;
  if (x > y) return 1;
  return 0;
}

bool RilsRols::dominates(const Fit& p, const Fit& f) { return p.e0 <= f.e0 && p.e1 <= f.e1 && p.size <= f.size; }
bool RilsRols::is_dominated(const std::vector<Fit>& par, const Fit& f) {
  for (const Fit& p : par) if (dominates(p, f)) return true;
  return false;
}
void RilsRols::add_to_pareto(std::vector<Fit>& par, const Fit& f) {   // 592-599
  if (is_dominated(par, f)) return;
  for (int i = (int)par.size() - 1; i >= 0; --i)
    if (dominates(f, par[i])) par.erase(par.begin() + i);
  par.push_back(f);
}

void RilsRols::print_state(const Fit& c) const {   // 563-579
  std::cout << "it=" << main_it_ << "\tfit_calls=" << fit_calls_ << "\tls_calls=" << ls_calls_;
  if (classification_) {
    std::cout << "\tcurr_LOSS=" << c.e0 << "\tcurr_size=" << c.size;
    std::cout << "\tfinal_LOSS=" << final_fitness_.e0 << "\tfinal_size=" << final_fitness_.size;
  } else {
    std::cout << "\tcurr_R2=" << (1 - c.e0) << "\tcurr_RMSE=" << c.e1 << "\tcurr_size=" << c.size;
    std::cout << "\tfinal_R2=" << (1 - final_fitness_.e0) << "\tfinal_RMSE=" << final_fitness_.e1
              << "\tfinal_size=" << final_fitness_.size;
  }
  std::cout << "\tchecks_skip=" << skipped_perts_ << "/" << total_perts_ << "\tsol=";
  std::cout << to_string(final_solution_) << std::endl << std::endl;
}

void RilsRols::log_move(int kind, const Flat& t, const Fit& f) {
  moves_.push_back(Move{kind, fit_calls_, f.e0, f.e1, f.size, to_string(t)});
}

// ------------------------------------------------------------------ generators
void RilsRols::setup_nodes(const std::vector<int>& rel) {   // 121-148
  const uint8_t ops[] = {PLUS, MINUS, MULTIPLY, DIVIDE, SIN, COS, LN, EXP, SQRT, SQR};
  for (uint8_t t : ops) allowed_.push_back(tok_op(t));
  const double consts[] = {-1., 0., 0.5, 1., 2., 3.14159265358979323846, 10.};
  for (double c : consts) allowed_.push_back(tok_const(c));
  for (int v : rel) allowed_.push_back(tok_var(v));
  if (classification_) {
    const uint8_t cops[] = {LT, GT, EQ, NEQ, MIN, MAX};
    for (uint8_t t : cops) allowed_.push_back(tok_op(t));
  }
  if (verbose_) std::cout << "Finished creating allowed nodes" << std::endl;
}

// node.cpp:97-118
static bool allowed_left(uint8_t parent, uint8_t child) {
  switch (parent) {
    case EXP: case LN: return !(child == EXP || child == LN);
    case POW: return child != POW;
    case COS: case SIN: return !(child == COS || child == SIN);
    default: return true;
  }
}

// get_all_subtrees (node.h:351-377): the subtree at 0, then BFS order
static void bfs_subtrees(const Flat& o, const std::vector<int32_t>& sz, std::vector<Flat>& out) {
  std::vector<int32_t> q{0};
  for (size_t h = 0; h < q.size(); ++h) {
    const int p = q[h];
    out.emplace_back(o.begin() + p, o.begin() + p + sz[p]);
    const int a = arity(o[p].t);
    if (a >= 1) q.push_back(p + 1);
    if (a >= 2) q.push_back(p + 1 + sz[p + 1]);
  }
}

static Flat join1(uint8_t op, const Flat& a) {
  Flat f;
  f.reserve(a.size() + 1);
  f.push_back(tok_op(op));
  f.insert(f.end(), a.begin(), a.end());
  return f;
}
static Flat join2(uint8_t op, const Flat& a, const Flat& b) {
  Flat f;
  f.reserve(a.size() + b.size() + 1);
  f.push_back(tok_op(op));
  f.insert(f.end(), a.begin(), a.end());
  f.insert(f.end(), b.begin(), b.end());
  return f;
}

void RilsRols::add_const_finetune(const Flat& o, std::vector<Flat>& c) {   // 163-179
  if (o[0].t != CONST) return;
  if (o[0].c == 0.0) { c.push_back({tok_const(-1.0)}); c.push_back({tok_const(1.0)}); return; }
  const double mult[] = {-1., 0.01, 0.1, 0.2, 0.5, 0.8, 0.9, 0., 1, 1.1, 1.2, 2., 3.14159265358979323846,
                         5., 10., 20., 50., 100.};
  for (double m : mult) c.push_back({tok_const(o[0].c * m)});
}

void RilsRols::add_change_to_subtree(const Flat& o, const std::vector<int32_t>& sz, std::vector<Flat>& c) {  // 200-209
  const int a = arity(o[0].t);
  if (a >= 1) {
    Flat l(o.begin() + 1, o.begin() + 1 + sz[1]);
    std::vector<int32_t> lsz(sz.begin() + 1, sz.begin() + 1 + sz[1]);
    bfs_subtrees(l, lsz, c);
  }
  if (a >= 2) {
    const int r = 1 + sz[1];
    Flat rt(o.begin() + r, o.begin() + r + sz[r]);
    std::vector<int32_t> rsz(sz.begin() + r, sz.begin() + r + sz[r]);
    bfs_subtrees(rt, rsz, c);
  }
}

void RilsRols::add_change_to_var_or_1(const Flat& o, std::vector<Flat>& c) {   // 222-232
  for (const Tok& n : allowed_) {
    if (n.t != VAR) continue;
    if (o[0].t == VAR && o[0].var == n.var) continue;
    c.push_back({n});
  }
  c.push_back({tok_const(1.0)});
}

void RilsRols::add_change_unary_applied(const Flat& o, std::vector<Flat>& c) {   // 245-252
  for (const Tok& n : allowed_) {
    if (arity(n.t) != 1 || !allowed_left(n.t, o[0].t)) continue;
    c.push_back(join1(n.t, o));
  }
}

void RilsRols::add_change_unary_to_another(const Flat& o, std::vector<Flat>& c) {   // 258-270
  if (arity(o[0].t) != 1) return;
  const Flat child(o.begin() + 1, o.end());
  for (const Tok& n : allowed_) {
    if (arity(n.t) != 1 || n.t == o[0].t) continue;
    if (!allowed_left(n.t, child[0].t)) continue;
    c.push_back(join1(n.t, child));
  }
}

void RilsRols::add_change_binary_applied(const Flat& o, const std::vector<int32_t>& sz, std::vector<Flat>& c) {  // 272-296
  std::vector<Flat> args;
  bfs_subtrees(o, sz, args);
  for (const Tok& n : allowed_)
    if (n.t == VAR || n.t == CONST) args.push_back({n});
  for (const Tok& nb : allowed_) {
    if (arity(nb.t) != 2) continue;
    for (const Flat& a : args) {
      if (allowed_left(nb.t, o[0].t)) c.push_back(join2(nb.t, o, a));
      if (!symmetric(nb.t) && allowed_left(nb.t, a[0].t)) c.push_back(join2(nb.t, a, o));
    }
  }
}

void RilsRols::add_change_binary_to_another(const Flat& o, const std::vector<int32_t>& sz, std::vector<Flat>& c) {  // 303-316
  if (arity(o[0].t) != 2) return;
  const Flat l(o.begin() + 1, o.begin() + 1 + sz[1]);
  const Flat r(o.begin() + 1 + sz[1], o.end());
  for (const Tok& nb : allowed_) {
    if (arity(nb.t) != 2 || nb.t == o[0].t) continue;
    if (allowed_left(nb.t, l[0].t)) c.push_back(join2(nb.t, l, r));
    if (!symmetric(nb.t) && allowed_left(nb.t, r[0].t)) c.push_back(join2(nb.t, r, l));
  }
}

void RilsRols::gen_perturb(const Flat& o, std::vector<Flat>& c) {   // 318-330
  std::vector<int32_t> sz;
  subtree_sizes(o, sz);
  add_change_to_subtree(o, sz, c);
  add_change_to_var_or_1(o, c);
  if (o[0].t == VAR) add_change_unary_applied(o, c);
  add_change_unary_to_another(o, c);
  if (o[0].t == VAR || o[0].t == CONST) add_change_binary_applied(o, sz, c);
  add_change_binary_to_another(o, sz, c);
}

void RilsRols::gen_change(const Flat& o, std::vector<Flat>& c) {   // 332-346
  std::vector<int32_t> sz;
  subtree_sizes(o, sz);
  add_const_finetune(o, c);
  add_change_to_subtree(o, sz, c);
  add_change_to_var_or_1(o, c);
  add_change_unary_applied(o, c);
  add_change_unary_to_another(o, c);
  add_change_binary_applied(o, sz, c);
  add_change_binary_to_another(o, sz, c);
}

// Which strings are the first of their value, in index order: open
// addressing over precomputed hashes, equal hashes settled by comparing the
// strings, so the outcome is that of inserting them into a set of strings.
// The table is at least twice the number of strings, so a probe ends.
class FirstSeen {
 public:
  FirstSeen(const std::vector<uint64_t>& hashes, const std::vector<std::string_view>& keys)
      : h_(hashes), k_(keys) {
    size_t cap = 16;
    while (cap < 2 * hashes.size()) cap <<= 1;
    slot_.assign(cap, -1);
    mask_ = cap - 1;
  }
  // true when no earlier insert had the string keys[i]
  bool insert(int32_t i) {
    for (size_t s = (size_t)h_[(size_t)i] & mask_;; s = (s + 1) & mask_) {
      const int32_t o = slot_[s];
      if (o < 0) {
        slot_[s] = i;
        return true;
      }
      if (h_[(size_t)o] == h_[(size_t)i] && k_[(size_t)o] == k_[(size_t)i]) return false;
    }
  }

 private:
  const std::vector<uint64_t>& h_;
  const std::vector<std::string_view>& k_;
  std::vector<int32_t> slot_;
  size_t mask_ = 0;
};

// rils_rols_cpp.cpp:348-443.  Candidate strings for the dedup are assembled
// from the solution's string around the changed site instead of re-printing
// the whole tree.  Three passes, the two dedups in the reference's order
// between them: the sites (every node, breadth first) generate their
// candidates and strings in parallel; the first of each string is built into
// a whole tree, rewritten and printed in parallel; the first of each
// rewritten string is returned.  Strings and trees live in per-site and
// per-thread buffers, so a pass allocates nothing per candidate.
std::vector<Flat> RilsRols::all_candidates(const Flat& sol, bool local_search) {
  std::vector<int32_t> sz;
  subtree_sizes(sol, sz);
  std::string full;
  std::vector<int32_t> be;
  append_string(sol, 0, full, &be);

  std::vector<int32_t> sites{0};   // every node, breadth first
  for (size_t h = 0; h < sites.size(); ++h) {
    const int p = sites[h];
    const int a = arity(sol[p].t);
    if (a >= 1) sites.push_back(p + 1);
    if (a >= 2) sites.push_back(p + 1 + sz[p + 1]);
  }
  struct Site {
    Flat toks;                      // the replacement subtrees, back to back
    std::vector<uint32_t> tend;     // end of each in toks
    std::string keys;               // the whole tree's string with each, back to back
    std::vector<uint32_t> kend;
    std::vector<uint64_t> hash;
  };
  std::vector<Site> per(sites.size());
  workers().parallel_for(sites.size(), [&](size_t i, int) {
    const int p = sites[i];
    const Flat old(sol.begin() + p, sol.begin() + p + sz[p]);
    std::vector<Flat> cands;
    if (local_search) gen_change(old, cands); else gen_perturb(old, cands);
    Site& S = per[i];
    S.tend.reserve(cands.size());
    S.kend.reserve(cands.size());
    S.hash.reserve(cands.size());
    S.keys.reserve(cands.size() * (full.size() + 16));
    for (const Flat& c : cands) {
      S.toks.insert(S.toks.end(), c.begin(), c.end());
      S.tend.push_back((uint32_t)S.toks.size());
      const size_t k0 = S.keys.size();
      S.keys.append(full, 0, be[2 * p]);
      append_string(c, 0, S.keys);
      S.keys.append(full, be[2 * p + 1], std::string::npos);
      S.kend.push_back((uint32_t)S.keys.size());
      S.hash.push_back(std::hash<std::string_view>{}(std::string_view(S.keys).substr(k0)));
    }
  }, 2, 1);

  struct Ref { int32_t site, j; };
  std::vector<std::string_view> keys;   // in the reference's order
  std::vector<uint64_t> hashes;
  std::vector<Ref> refs;
  for (size_t i = 0; i < per.size(); ++i) {
    const Site& S = per[i];
    for (size_t j = 0; j < S.kend.size(); ++j) {
      const uint32_t k0 = j ? S.kend[j - 1] : 0;
      keys.emplace_back(S.keys.data() + k0, S.kend[j] - k0);
      hashes.push_back(S.hash[j]);
      refs.push_back(Ref{(int32_t)i, (int32_t)j});
    }
  }
  std::vector<Ref> kept;
  kept.reserve(refs.size());
  {
    FirstSeen first(hashes, keys);
    for (size_t i = 0; i < refs.size(); ++i)
      if (first.insert((int32_t)i)) kept.push_back(refs[i]);
  }

  struct Done { int32_t tid; uint32_t t0, t1, s0, s1; uint64_t hash; };
  std::vector<Done> done(kept.size());
  for (GenScratch& g : gen_) { g.toks.clear(); g.strs.clear(); }
  workers().parallel_for(kept.size(), [&](size_t i, int tid) {
    const Site& S = per[(size_t)kept[i].site];
    const int p = sites[(size_t)kept[i].site];
    const int j = kept[i].j;
    GenScratch& g = gen_[(size_t)tid];
    Flat& f = g.tree;
    f.clear();
    f.insert(f.end(), sol.begin(), sol.begin() + p);
    f.insert(f.end(), S.toks.begin() + (j ? S.tend[j - 1] : 0), S.toks.begin() + S.tend[j]);
    f.insert(f.end(), sol.begin() + p + sz[p], sol.end());
    rewrite_rounds(f, true, pools_[(size_t)tid]);
    Done& d = done[i];
    d.tid = tid;
    d.t0 = (uint32_t)g.toks.size();
    g.toks.insert(g.toks.end(), f.begin(), f.end());
    d.t1 = (uint32_t)g.toks.size();
    d.s0 = (uint32_t)g.strs.size();
    append_string(f, 0, g.strs);
    d.s1 = (uint32_t)g.strs.size();
    d.hash = std::hash<std::string_view>{}(std::string_view(g.strs).substr(d.s0));
  });

  std::vector<std::string_view> strs(done.size());
  std::vector<uint64_t> h2(done.size());
  for (size_t i = 0; i < done.size(); ++i) {
    const Done& d = done[i];
    strs[i] = std::string_view(gen_[(size_t)d.tid].strs.data() + d.s0, d.s1 - d.s0);
    h2[i] = d.hash;
  }
  FirstSeen first2(h2, strs);
  std::vector<Flat> out;
  out.reserve(done.size());
  for (size_t i = 0; i < done.size(); ++i) {
    if (!first2.insert((int32_t)i)) continue;
    const Flat& t = gen_[(size_t)done[i].tid].toks;
    out.emplace_back(t.begin() + done[i].t0, t.begin() + done[i].t1);
  }
  return out;
}

// ------------------------------------------------------------------ fitness
static Fit to_fit(const rr_fit& r, int size) {
  if (r.flags & RR_F_SENTINEL) return Fit{1000, 1000, 1000};   // rils_rols_cpp.cpp:536-537
  return Fit{r.e0, r.e1, size};
}

Fit RilsRols::fitness_raw(const Flat& t) {   // fitness(), 520-541
  fit_calls_++;
  std::vector<const Flat*> one{&t};
  std::vector<rr_fit> out;
  gpu_->raw(one, out, stats_);
  stats_.single_calls++;
  return to_fit(out[0], (int)t.size());
}

Flat RilsRols::tune_and_fit(const Flat& cand, Fit* fit) {   // tune_constants (445-518) + fitness
  std::vector<std::vector<Flat>> facs(1);
  ols_factors(cand, facs[0], pool_);
  std::vector<double> coefs;
  std::vector<int64_t> off;
  std::vector<rr_fit> out;
  gpu_->ols(facs, 1, RR_FP64, coefs, off, out, stats_);
  stats_.single_calls++;
  Flat tree;
  ols_tree(facs[0], coefs.data(), tree);
  fit_calls_++;
  *fit = to_fit(out[0], (int)tree.size());
  return tree;
}

// FP32 sweeps.  A certified FP32 fitness f32 with relative bound err puts the
// FP64 fitness value in [f32 (1 - err), f32 (1 + err)]; the OLS-tree size is
// exact (the kernel certifies the coefficient threshold decisions).  A
// candidate is rejected on FP32 evidence only when compare_fitness(tf, curr)
// < 0 is impossible for every fitness in that interval; every other candidate
// (possible acceptance, uncertified, or near the incumbent) takes its FP64
// result, so acceptance decisions, the accepted trees and their coefficients
// are those of the FP64 form.
bool RilsRols::fp32_rejects(const rr_fit& r, const std::vector<int32_t>& fsz, const double* c,
                            const Fit& curr) const {
  if (r.flags & RR_F_NEEDS64) return false;
  const int size = (r.flags & RR_F_SENTINEL) ? 1000 : ols_tree_size(fsz, c);
  if ((size > max_complexity_ || curr.size > max_complexity_) && size != curr.size) return size - curr.size > 0;
  const Fit tf = to_fit(r, size);
  const double lo = fitness_value(tf) * (1.0 - r.err);
  return lo >= fitness_value(curr);
}

// One FP64 batch, before the replay, for the candidates the replay is likely
// to need: those that could beat the incumbent the sweep starts from (the
// sweep's RR_FP32_CHECKED call already returned FP64 results, RR_F_FP64, for
// every candidate the FP32 form could not certify).  Its results replace the
// FP32 ones in place.
void RilsRols::rescore_speculative(const std::vector<std::vector<Flat>>& facs, size_t m, const Fit& curr,
                                   std::vector<double>& coefs, const std::vector<int64_t>& off,
                                   std::vector<rr_fit>& res, std::vector<uint8_t>& is64) {
  is64.assign(m, 0);
  std::vector<int32_t> idx;
  std::vector<int32_t> fsz;
  for (size_t j = 0; j < m; ++j) {
    if (res[j].flags & RR_F_FP64) {
      res[j].flags &= ~RR_F_FP64;
      is64[j] = 1;
      stats_.rescored_fp64++;
      continue;
    }
    fsz.clear();
    for (const Flat& f : facs[j]) fsz.push_back((int32_t)f.size());
    if (!fp32_rejects(res[j], fsz, coefs.data() + off[j], curr)) idx.push_back((int32_t)j);
  }
  if (idx.empty()) return;
  std::vector<double> c64;
  std::vector<int64_t> o64;
  std::vector<rr_fit> r64;
  gpu_->ols_subset(facs, idx.data(), idx.size(), RR_FP64, c64, o64, r64, stats_);
  for (size_t q = 0; q < idx.size(); ++q) {
    const size_t j = (size_t)idx[q];
    std::copy(c64.begin() + o64[q], c64.begin() + o64[q] + (int64_t)facs[j].size() + 1, coefs.begin() + off[j]);
    res[j] = r64[q];
    is64[j] = 1;
  }
  stats_.rescored_fp64 += (int64_t)idx.size();
}

// Make res[j] and the coefficients at coefs[off[j]] the FP64 results of
// candidate j, scoring it alone unless the speculative batch already did.
void RilsRols::ensure_fp64(const std::vector<std::vector<Flat>>& facs, size_t j, std::vector<double>& coefs,
                           const std::vector<int64_t>& off, std::vector<rr_fit>& res, std::vector<uint8_t>& is64) {
  if (is64[j]) return;
  std::vector<double> c64;
  std::vector<int64_t> o64;
  std::vector<rr_fit> r64;
  const int32_t one = (int32_t)j;
  gpu_->ols_subset(facs, &one, 1, RR_FP64, c64, o64, r64, stats_);
  std::copy(c64.begin(), c64.begin() + (int64_t)facs[j].size() + 1, coefs.begin() + off[j]);
  res[j] = r64[0];
  is64[j] = 1;
  stats_.rescored_fp64++;
  stats_.single_calls++;
}

Flat RilsRols::local_search(const Flat& passed, Fit* out_fit) {   // 601-643
  std::vector<Fit> pareto;
  ls_calls_++;
  bool improved = true;
  Fit curr_fit;
  Flat curr = tune_and_fit(passed, &curr_fit);
  std::vector<std::vector<Flat>> facs;
  std::vector<double> coefs;
  std::vector<int64_t> off;
  std::vector<rr_fit> res;
  std::vector<int32_t> fsz;
  std::vector<uint8_t> is64;
  while (improved && !finished()) {
    improved = false;
    const auto sweep0 = std::chrono::steady_clock::now();
    auto t0 = sweep0;
    std::vector<Flat> perts = all_candidates(curr, true);
    stats_.host_gen_s += secs_since(t0);

    // every candidate the replay can reach costs at least one fit call
    const int64_t room = std::max<int64_t>((int64_t)max_fit_calls_ - fit_calls_, 0);
    const size_t m = std::min<size_t>(perts.size(), (size_t)room);
    t0 = std::chrono::steady_clock::now();
    if (facs.size() < m) facs.resize(m);
    workers().parallel_for(m, [&](size_t j, int tid) { ols_factors(perts[j], facs[j], pools_[(size_t)tid]); });
    stats_.host_prep_s += secs_since(t0);
    const bool fp32 = precision_ == RR_FP32_REFINED;
    if (autotune_ && !tuned_ && m >= 256) {
      gpu_->tune_next = true;
      tuned_ = true;
    }
    gpu_->ols(facs, m, fp32 ? RR_FP32_CHECKED : RR_FP64, coefs, off, res, stats_);
    stats_.sweeps++;
    stats_.batched_cands += (int64_t)m;
    if (fp32) rescore_speculative(facs, m, curr_fit, coefs, off, res, is64);

    t0 = std::chrono::steady_clock::now();
    for (size_t j = 0; j < perts.size(); ++j) {
      if (finished()) break;
      if (j >= m) throw std::logic_error("local-search replay ran past the scored batch");
      const double* cj = coefs.data() + off[j];
      fsz.clear();
      for (const Flat& f : facs[j]) fsz.push_back((int32_t)f.size());
      fit_calls_++;
      Fit tf;
      if (fp32 && !is64[j]) {
        // an FP32 result only ever rejects; anything that might be accepted is decided in FP64
        if (fp32_rejects(res[j], fsz, cj, curr_fit)) continue;
        ensure_fp64(facs, j, coefs, off, res, is64);
      }
      tf = to_fit(res[j], ols_tree_size(fsz, cj));
      if (verbose_ && fit_calls_ % 10000 == 0) print_state(curr_fit);
      if (!is_dominated(pareto, tf) && compare_fitness(tf, curr_fit) < 0) {
        improved = true;
        Flat tuned;
        ols_tree(facs[j], cj, tuned);
        rewrite_rounds(tuned, false, pool_);   // expand + simplify rounds, 621-630
        tf = fitness_raw(tuned);
        curr = tuned;
        curr_fit = tf;
        add_to_pareto(pareto, tf);
        log_move(0, curr, curr_fit);
      }
    }
    stats_.replay_s += secs_since(t0);
    sweep_ms_.push_back((float)(1e3 * secs_since(sweep0)));
  }
  *out_fit = curr_fit;
  return curr;
}

// canonical-order sum of the FP64 form (rrm::canon_sum)
static double canon_sum_host(const double* v, int64_t n, const double* w = nullptr, double wshift = 0.0) {
  return rrm::canon_sum(0, n, [&](int64_t i) {
    return w ? (v[i] - w[i]) * (v[i] - w[i]) : (v[i] - wshift) * (v[i] - wshift);
  });
}

// 753-770: with more than 200 features keep the 200 with the largest R2(X[i], y)
std::vector<int> RilsRols::relevant_features(const std::vector<std::vector<double>>& X, const std::vector<double>& y) {
  const int d = (int)X.size();
  std::vector<int> rel;
  if (d <= max_feat_) { rel.resize(d); std::iota(rel.begin(), rel.end(), 0); return rel; }
  const int64_t n = (int64_t)y.size();
  std::vector<std::tuple<double, int>> by;
  for (int i = 0; i < d; ++i) {
    const double* x = X[i].data();
    const double xbar = rrm::canon_sum(0, n, [&](int64_t r) { return x[r]; }) / (double)n;
    const double ssr = canon_sum_host(x, n, y.data());
    const double sst = canon_sum_host(x, n, nullptr, xbar);
    by.push_back(std::tuple<double, int>{1 - ssr / sst, i});
  }
  std::sort(by.begin(), by.end(), std::greater<>());
  for (int i = 0; i < max_feat_; ++i) rel.push_back(std::get<1>(by[i]));
  return rel;
}

void RilsRols::fit_inner(const std::vector<std::vector<double>>& X_all, const std::vector<double>& y_all) {  // 772-859
  reset();
  moves_.clear();
  stats_ = Stats();
  sweep_ms_.clear();
  tuned_ = false;
  const auto t_setup = std::chrono::steady_clock::now();
  const int64_t N = (int64_t)y_all.size();
  const int sample_cnt = int(sample_size_ * (double)N);
  std::vector<int> selected(N);
  std::iota(selected.begin(), selected.end(), 0);
  std::shuffle(selected.begin(), selected.end(), std::default_random_engine(random_state_));
  std::vector<std::vector<double>> X(X_all.size(), std::vector<double>(std::max(sample_cnt, 0)));
  std::vector<double> y(std::max(sample_cnt, 0));
  for (int ix = 0; ix < sample_cnt; ++ix) {
    const int i = selected[ix];
    for (size_t j = 0; j < X.size(); ++j) X[j][ix] = X_all[j][i];
    y[ix] = y_all[i];
  }
  if (sample_cnt <= 0) throw std::invalid_argument("sample_size * n must be at least one sample");
  std::vector<int> rel = relevant_features(X, y);
  // The reference appends to its node list on every fit of one object (its
  // reset() leaves the list alone).  The list is only iterated by the
  // candidate generators, and every candidate passes the first-occurrence
  // string dedup, so the copies added only repeat candidates that dedup
  // drops: clearing it gives the same candidates in the same order, without
  // the growing generation time of later fits.
  allowed_.clear();
  setup_nodes(rel);
  if (!gpu_) gpu_ = new GpuSession(backend_, device_, slab_budget_mb_, ctx_options_);
  gpu_->upload(X, y);
  stats_.setup_s = secs_since(t_setup);

  final_solution_ = {tok_const(0.0)};
  final_fitness_ = fitness_raw(final_solution_);
  bool improved = true;
  std::vector<rr_fit> res;
  while (!finished()) {
    main_it_ += 1;
    auto tm = std::chrono::steady_clock::now();
    Flat start = final_solution_;
    if (!improved) {
      std::vector<Flat> all_perts = all_candidates(final_solution_, false);
      const Flat p1 = all_perts[(size_t)rng_.next() % all_perts.size()];
      std::vector<Flat> all2 = all_candidates(p1, false);
      start = all2[(size_t)rng_.next() % all2.size()];
      if (verbose_) std::cout << "Randomized to " << to_string(start) << std::endl;
    }
    improved = false;
    std::vector<Flat> all_perts = all_candidates(start, false);
    if (verbose_) std::cout << "Checking " << all_perts.size() << " perturbations of starting solution." << std::endl;
    // 819-830: the perturbations that get scored, then one batched score
    std::vector<const Flat*> sel;
    {
      const auto el = [&]() {
        return std::chrono::duration_cast<std::chrono::seconds>(std::chrono::steady_clock::now() - start_time_).count();
      };
      int64_t virt = fit_calls_;
      for (size_t i = 0; i < all_perts.size(); ++i) {
        if (virt >= max_fit_calls_ || el() > max_seconds_) break;
        if (check_skip(to_string(all_perts[i]))) continue;
        sel.push_back(&all_perts[i]);
        ++virt;
      }
    }
    stats_.host_main_s += secs_since(tm);
    gpu_->raw(sel, res, stats_);
    fit_calls_ += (int)sel.size();
    tm = std::chrono::steady_clock::now();
    std::vector<std::pair<double, const Flat*>> by;
    by.reserve(sel.size());
    for (size_t i = 0; i < sel.size(); ++i) by.emplace_back(to_fit(res[i], (int)sel[i]->size()).e0, sel[i]);
    std::sort(by.begin(), by.end(), [](const std::pair<double, const Flat*>& a, const std::pair<double, const Flat*>& b) {
      return a.first < b.first;
    });
    stats_.host_main_s += secs_since(tm);
    for (size_t i = 0; i < by.size(); ++i) {   // 833-855
      if (finished()) break;
      const Flat& lp = *by[i].second;
      checked_perts_.insert(to_string(lp));
      Fit ls_fit;
      Flat ls = local_search(lp, &ls_fit);
      fit_calls_++;   // fitness(ls_pert), 842: the same tree local_search just scored
      if (compare_fitness(ls_fit, final_fitness_) < 0) {
        improved = true;
        final_solution_ = ls;
        final_fitness_ = ls_fit;
        if (verbose_) print_state(final_fitness_);
        best_time_ = std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - start_time_).count() / 1000.0;
        log_move(1, final_solution_, final_fitness_);
      }
    }
  }
  total_time_ = std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - start_time_).count() / 1000.0;
}

void RilsRols::fit(const double* Xp, int64_t x_size, const double* yp, int64_t y_size, int data_cnt, int feat_cnt) {
  if (x_size != (int64_t)data_cnt * feat_cnt)
    throw std::invalid_argument("Size of X " + std::to_string(x_size) +
                                " is not the same as the product of data count and feature count " +
                                std::to_string((int64_t)data_cnt * feat_cnt));
  if (y_size != data_cnt)
    throw std::invalid_argument("Size of y " + std::to_string(y_size) + " is not the same as the data count " +
                                std::to_string(data_cnt));
  std::vector<std::vector<double>> Xe(feat_cnt, std::vector<double>(data_cnt));
  for (int i = 0; i < data_cnt; ++i)
    for (int j = 0; j < feat_cnt; ++j) Xe[j][i] = Xp[(int64_t)i * feat_cnt + j];
  std::vector<double> ye(yp, yp + data_cnt);
  fit_inner(Xe, ye);
}

void RilsRols::local_search_from(const double* Xp, int64_t x_size, const double* yp, int64_t y_size, int data_cnt,
                                 int feat_cnt, const std::string& prefix) {
  if (x_size != (int64_t)data_cnt * feat_cnt || y_size != data_cnt || data_cnt <= 0)
    throw std::invalid_argument("X must be data_cnt x feat_cnt and y of length data_cnt");
  const Flat start = parse_prefix(prefix);
  std::vector<std::vector<double>> X(feat_cnt, std::vector<double>(data_cnt));
  for (int i = 0; i < data_cnt; ++i)
    for (int j = 0; j < feat_cnt; ++j) X[j][i] = Xp[(int64_t)i * feat_cnt + j];
  std::vector<double> y(yp, yp + data_cnt);
  reset();
  moves_.clear();
  stats_ = Stats();
  sweep_ms_.clear();
  allowed_.clear();
  std::vector<int> rel(feat_cnt);
  std::iota(rel.begin(), rel.end(), 0);
  setup_nodes(rel);
  if (!gpu_) gpu_ = new GpuSession(backend_, device_, slab_budget_mb_, ctx_options_);
  gpu_->upload(X, y);
  Fit f;
  final_solution_ = local_search(start, &f);
  final_fitness_ = f;
}

std::vector<double> RilsRols::predict(const double* Xp, int64_t x_size, int data_cnt, int feat_cnt) {
  if (x_size != (int64_t)data_cnt * feat_cnt)
    throw std::invalid_argument("Size of X " + std::to_string(x_size) +
                                " is not the same as the product of data count and feature count " +
                                std::to_string((int64_t)data_cnt * feat_cnt));
  if (final_solution_.empty()) throw std::runtime_error("predict called before fit");
  if (!gpu_) gpu_ = new GpuSession(backend_, device_, slab_budget_mb_, ctx_options_);
  rr_ctx* ctx = gpu_->be->context();
  if (!ctx) throw std::runtime_error("predict needs a fitness backend with a device context");
  std::vector<double> fm((size_t)data_cnt * std::max(feat_cnt, 1));
  for (int i = 0; i < data_cnt; ++i)
    for (int j = 0; j < feat_cnt; ++j) fm[(size_t)j * data_cnt + i] = Xp[(int64_t)i * feat_cnt + j];
  std::vector<double> zeros(data_cnt, 0.0), out(data_cnt);
  int32_t h = -1;
  rr_check(rr_dataset_upload(ctx, fm.data(), zeros.data(), data_cnt, feat_cnt, &h), "rr_dataset_upload");
  std::vector<uint32_t> code;
  std::vector<double> consts;
  compile(final_solution_, code, consts);
  const int rc = rr_predict(ctx, h, code.data(), (int32_t)code.size(), consts.data(), (int32_t)consts.size(),
                            out.data());
  rr_dataset_free(ctx, h);
  rr_check(rc, "rr_predict");
  if (classification_)
    for (double& v : out) v = v >= 0.5 ? 1.0 : 0.0;
  return out;
}

std::string RilsRols::get_model_string() const {
  if (final_solution_.empty()) throw std::runtime_error("no model: call fit first");
  return to_string(final_solution_);
}

}  // namespace rrh
