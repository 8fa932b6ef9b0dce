// search.h -- the RILS-ROLS search driver of the product path.
//
// Mirrors class rils_rols of the reference (rils_rols_cpp/rils_rols_cpp.cpp:
// 88-881): the same parameters, counters, candidate generators, acceptance
// rules and budget.  What changes is where fitness is computed: every
// local-search sweep (rils_rols_cpp.cpp:610-639) is scored by one batched call
// through the C ABI of include/rr_fitness.h, and the host then replays the
// sweep's acceptance decisions in order, so moves, fit_calls and the model are
// those of the sequential loop.
#pragma once
#include <stdint.h>
#include <stdlib.h>

#include <chrono>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../include/rr_fitness.h"
#include "tree.h"
#include "workers.h"

namespace rrh {

struct Fit {
  double e0, e1;
  int size;
};

struct Move {
  int kind;          // 0 local-search acceptance, 1 new best solution
  int fit_calls;
  double e0, e1;
  int size;
  std::string model;
};

struct Stats {
  int64_t sweeps = 0, batched_cands = 0, raw_batches = 0, raw_cands = 0, single_calls = 0;
  int64_t rescored_fp64 = 0;
  double host_gen_s = 0, host_prep_s = 0, gpu_s = 0, kernel_ms = 0, replay_s = 0;
  double host_main_s = 0;   // the main loop's perturbation lists, skip checks and sort
  double setup_s = 0;       // sampling, feature filter, device context and dataset upload
};

// glibc's rand() sequence with per-object state (random_r on a TYPE_3 state,
// the generator rand() itself uses), so concurrent searches stay independent.
class CRand {
 public:
  void seed(unsigned s);
  int next();
 private:
  char state_[128];
  struct random_data data_ {};
};

// Where a search's fitness calls go.  The default is a context of its own
// (DirectBackend in search.cpp); the request loop of serve.h passes a shared
// dispatcher that merges the calls of many concurrent searches into one
// launch.  Implementations are thread-safe where they are shared.
class FitnessBackend {
 public:
  virtual ~FitnessBackend() = default;
  // upload a dataset (feature-major X, n x d); returns its handle
  virtual int32_t upload(const double* X_fm, const double* y, int64_t n, int32_t d) = 0;
  virtual void release(int32_t handle) = 0;
  // one batched fitness call (rr_fitness_batch semantics); kernel_ms gets the
  // kernel time of the launch that carried it
  virtual void fitness(int32_t handle, const rr_batch& b, int32_t mode, int32_t precision, double* coefs,
                       rr_fit* fit, double* kernel_ms) = 0;
  // a context for predict, or null when the backend has none to lend
  virtual rr_ctx* context() = 0;
  // pick the kernel schedule on a representative batch (no-op by default)
  virtual void autotune(int32_t handle, const rr_batch& b, int32_t mode, int32_t precision) {
    (void)handle; (void)b; (void)mode; (void)precision;
  }
};

class GpuSession;

class RilsRols {
 public:
  RilsRols(bool classification, int max_fit_calls, int max_seconds, double complexity_penalty, int max_complexity,
           double sample_size, bool verbose, int random_state);
  ~RilsRols();

  // rils_rols_cpp.cpp:717-728 (X row-major, data_cnt x feat_cnt)
  void fit(const double* X, int64_t x_size, const double* y, int64_t y_size, int data_cnt, int feat_cnt);
  // One local_search (rils_rols_cpp.cpp:601-643) from the tree in `prefix`
  // (exchange format) on X row-major, every feature allowed, no sampling;
  // the result becomes the model.  Beyond the reference's API: it lets a
  // caller (and the tests) drive the batched sweep without the main loop.
  void local_search_from(const double* X, int64_t x_size, const double* y, int64_t y_size, int data_cnt,
                         int feat_cnt, const std::string& prefix);
  // rils_rols_cpp.cpp:730-750
  std::vector<double> predict(const double* X, int64_t x_size, int data_cnt, int feat_cnt);

  std::string get_model_string() const;
  double get_best_time() const { return best_time_; }
  double get_total_time() const { return total_time_; }
  int get_fit_calls() const { return fit_calls_; }

  // beyond the reference: move log, timing, precision of the batched sweeps
  const std::vector<Move>& moves() const { return moves_; }
  const Stats& stats() const { return stats_; }
  void set_precision(int p) { precision_ = p; }
  int precision() const { return precision_; }
  void set_device(int d) { device_ = d; }
  void set_slab_budget_mb(int64_t mb) { slab_budget_mb_ = mb; }
  // an rr_set_option for the context the search creates (its own backend
  // only; applied when the context is made, in call order)
  void set_context_option(const std::string& name, int64_t value) { ctx_options_.emplace_back(name, value); }
  void set_host_threads(int t);
  // route fitness calls through `be` (not owned; must outlive the fit)
  void set_backend(FitnessBackend* be) { backend_ = be; }
  // time the first sweep of at least 256 candidates under every kernel
  // schedule and keep the fastest (rr_autotune); off by default
  void set_autotune(bool on) { autotune_ = on; }
  // per local-search sweep: wall time of batch + replay, ms
  const std::vector<float>& sweep_ms() const { return sweep_ms_; }
  std::string final_prefix() const { return final_solution_.empty() ? "" : to_prefix(final_solution_); }

  // generators (rils_rols_cpp.cpp:121-443) and the feature filter
  // (753-770), public for tests
  void setup_nodes(const std::vector<int>& rel_feat);
  std::vector<int> relevant_features(const std::vector<std::vector<double>>& X, const std::vector<double>& y);
  std::vector<Flat> all_candidates(const Flat& solution, bool local_search);

 private:
  // parameters (rils_rols_cpp.cpp:92-94)
  int max_fit_calls_, max_seconds_, random_state_, max_feat_ = 200;
  double complexity_penalty_, sample_size_, max_complexity_;
  bool classification_, verbose_;
  // state (97-104)
  int main_it_ = 0, last_improved_it_ = 0, fit_calls_ = 0, ls_calls_ = 0, skipped_perts_ = 0, total_perts_ = 0;
  std::unordered_set<std::string> checked_perts_;
  std::chrono::steady_clock::time_point start_time_;
  Flat final_solution_;
  Fit final_fitness_{0, 0, 0};
  double best_time_ = 0, total_time_ = 0;
  std::vector<Tok> allowed_;
  CRand rng_;
  std::vector<Move> moves_;
  Stats stats_;
  int precision_ = RR_FP64;
  int device_ = 0;
  int64_t slab_budget_mb_ = 0;
  std::vector<std::pair<std::string, int64_t>> ctx_options_;
  GpuSession* gpu_ = nullptr;
  FitnessBackend* backend_ = nullptr;
  bool autotune_ = false, tuned_ = false;
  std::vector<float> sweep_ms_;
  Pool pool_;
  int host_threads_ = 0;                  // 0: min(hardware threads, 8)
  Workers* workers_ = nullptr;
  std::vector<Pool> pools_;               // one per worker thread
  struct alignas(64) GenScratch {         // per worker thread: all_candidates' reusable buffers
    Flat tree;                            // the candidate being rewritten
    Flat toks;                            // rewritten candidates, back to back
    std::string strs;                     // their strings, back to back
  };
  std::vector<GenScratch> gen_;           // one per worker thread
  Workers& workers();

  void reset();
  bool finished() const;
  bool check_skip(const std::string& s);
  double fitness_value(const Fit& f) const;
  int compare_fitness(const Fit& a, const Fit& b) const;
  static bool dominates(const Fit& p, const Fit& f);
  static bool is_dominated(const std::vector<Fit>& pareto, const Fit& f);
  // FP32 sweeps (search.cpp)
  bool fp32_rejects(const rr_fit& r, const std::vector<int32_t>& fsz, const double* c, const Fit& curr) const;
  void rescore_speculative(const std::vector<std::vector<Flat>>& facs, size_t m, const Fit& curr,
                           std::vector<double>& coefs, const std::vector<int64_t>& off, std::vector<rr_fit>& res,
                           std::vector<uint8_t>& is64);
  void ensure_fp64(const std::vector<std::vector<Flat>>& facs, size_t j, std::vector<double>& coefs,
                   const std::vector<int64_t>& off, std::vector<rr_fit>& res, std::vector<uint8_t>& is64);
  static void add_to_pareto(std::vector<Fit>& pareto, const Fit& f);
  void print_state(const Fit& curr) const;
  void log_move(int kind, const Flat& t, const Fit& f);

  // candidate generators on flat subtrees
  void gen_change(const Flat& old, std::vector<Flat>& out);
  void gen_perturb(const Flat& old, std::vector<Flat>& out);
  void add_const_finetune(const Flat& o, std::vector<Flat>& c);
  void add_change_to_subtree(const Flat& o, const std::vector<int32_t>& sz, std::vector<Flat>& c);
  void add_change_to_var_or_1(const Flat& o, std::vector<Flat>& c);
  void add_change_unary_applied(const Flat& o, std::vector<Flat>& c);
  void add_change_unary_to_another(const Flat& o, std::vector<Flat>& c);
  void add_change_binary_applied(const Flat& o, const std::vector<int32_t>& sz, std::vector<Flat>& c);
  void add_change_binary_to_another(const Flat& o, const std::vector<int32_t>& sz, std::vector<Flat>& c);

  // fitness through the GPU
  Fit fitness_raw(const Flat& t);                     // one fit call
  Flat tune_and_fit(const Flat& cand, Fit* fit);      // tune_constants + fitness, one fit call
  Flat local_search(const Flat& passed, Fit* out_fit);
  void fit_inner(const std::vector<std::vector<double>>& X_all, const std::vector<double>& y_all);
};

}  // namespace rrh
