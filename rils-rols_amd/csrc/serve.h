// serve.h -- request loop that keeps many independent RILS-ROLS runs resident
// on one GPU and batches their fitness work as it arrives.
//
// Every admitted run executes the product search (search.h) on a host thread
// of its own.  Its fitness calls do not go to a device context of its own but
// to the shared Dispatcher: one thread that owns the device context, takes
// every request pending at that moment (continuous batching: requests that
// arrive while a launch is on the GPU join the next one), merges those of the
// same mode and precision into ONE rr_fitness_multi launch over all their
// datasets, and hands each run its slice of the results.  Per candidate the
// kernels' arithmetic does not depend on what else is in the launch, so a run
// produces the same moves and model alone and in the mix.
//
// Admission: a run is admitted when fewer than max_active runs are resident
// and its HBM footprint (its dataset: FP64 X and y, the FP32 copy of X and a
// column of ones, at the padded row stride) fits in the budget beside the resident ones;
// otherwise it waits in FIFO order.  The work slab of the FP64 kernels is
// shared and bounded separately (rr_set_option "slab_budget_mb").
#pragma once
#include <stdint.h>

#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rr_fitness.h"
#include "search.h"

namespace rrh {

// Merge several batches into one (constant pools concatenated, constant
// operands of the code renumbered).  Exposed for tests.
struct BatchView {
  const rr_batch* b;
  int32_t dataset;
};
void merge_batches(const std::vector<BatchView>& parts, std::vector<int32_t>& cand_start,
                   std::vector<int32_t>& prog_start, std::vector<uint32_t>& code, std::vector<double>& consts,
                   std::vector<int32_t>& handles);

class Dispatcher : public FitnessBackend {
 public:
  // gather_us: when fewer fitness requests are pending than runs hold a
  // resident dataset, wait up to this long for more before launching (0:
  // launch whatever is pending at once)
  Dispatcher(int device, int64_t slab_budget_mb, int gather_us = 0);
  ~Dispatcher() override;
  int32_t upload(const double* X_fm, const double* y, int64_t n, int32_t d) override;
  void release(int32_t handle) override;
  void fitness(int32_t handle, const rr_batch& b, int32_t mode, int32_t precision, double* coefs, rr_fit* fit,
               double* kernel_ms) override;
  rr_ctx* context() override { return nullptr; }

  struct Counters {
    int64_t launches = 0, requests = 0, candidates = 0, max_requests_per_launch = 0;
    double kernel_ms = 0, busy_s = 0, gather_s = 0;
  };
  Counters counters();

 private:
  struct Req;
  struct Flight;
  void loop();
  bool merge(Flight& f);
  void finish(Flight& f, int rc, double ms);
  void retire_oldest();
  rr_ctx* ctx_ = nullptr;
  std::mutex m_;
  std::condition_variable cv_;
  std::deque<Req*> q_;
  bool stop_ = false;
  std::string init_error_;
  bool ready_ = false;
  int gather_us_ = 0;
  int live_ = 0;   // resident datasets: one per run that is searching
  Counters ctr_;
  std::deque<std::unique_ptr<Flight>> flights_;   // submitted launches not yet waited for (dispatcher thread only)
  std::thread th_;
};

struct RunSpec {
  std::vector<double> X;   // row-major n x d
  std::vector<double> y;
  int n = 0, d = 0;
  bool classification = false;
  int max_fit_calls = 100000, max_seconds = 100;
  double complexity_penalty = 0.001;
  int max_complexity = 50;
  double sample_size = 1.0;
  int random_state = 0;
  int precision = RR_FP64;
};

struct RunResult {
  int id = -1;
  bool ok = false;
  std::string error, model, prefix;
  int fit_calls = 0;
  std::vector<Move> moves;
  std::vector<float> sweep_ms;
  Stats stats;
  double submitted_s = 0, admitted_s = 0, finished_s = 0;   // since server start
  int64_t footprint = 0;
};

class Server {
 public:
  Server(int device, int64_t hbm_budget_mb, int max_active, int host_threads_per_run, int64_t slab_budget_mb,
         int gather_us = 0);
  ~Server();
  int submit(RunSpec spec);
  bool done(int id);
  RunResult wait(int id);
  Dispatcher::Counters dispatcher_counters() { return disp_->counters(); }
  int64_t resident_bytes();
  static int64_t footprint(int64_t n_samples, int d);

 private:
  struct Run;
  void admit_locked();
  void run_body(Run* r);
  double now() const;
  std::unique_ptr<Dispatcher> disp_;
  int64_t budget_;
  int max_active_, threads_per_run_;
  std::mutex m_;
  std::condition_variable cv_;
  std::map<int, std::unique_ptr<Run>> runs_;
  std::deque<int> pending_;
  int active_ = 0, next_id_ = 0;
  int64_t used_ = 0;
  std::chrono::steady_clock::time_point t0_;
};

}  // namespace rrh
