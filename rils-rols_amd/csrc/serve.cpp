// serve.cpp -- request loop of many resident RILS-ROLS runs (see serve.h).
#include "serve.h"

#include <algorithm>
#include <memory>
#include <stdexcept>

namespace rrh {

// ------------------------------------------------------------------ merging
static bool uses_const(uint32_t w) {
  const uint32_t op = w & 0xffu;
  if (op == RR_OP_LDC) return true;
  return op >= RR_OP_BIN && ((op - RR_OP_BIN) & 7u) >= RR_M_C;
}

void merge_batches(const std::vector<BatchView>& parts, std::vector<int32_t>& cand_start,
                   std::vector<int32_t>& prog_start, std::vector<uint32_t>& code, std::vector<double>& consts,
                   std::vector<int32_t>& handles) {
  cand_start.assign(1, 0);
  prog_start.assign(1, 0);
  code.clear();
  consts.clear();
  handles.clear();
  for (const BatchView& pv : parts) {
    const rr_batch& b = *pv.b;
    const uint32_t kofs = (uint32_t)consts.size();
    if (b.n_consts > 0) consts.insert(consts.end(), b.consts, b.consts + b.n_consts);
    const int32_t pbase = (int32_t)prog_start.size() - 1;
    const int32_t np = b.n_cands > 0 ? b.cand_start[b.n_cands] : 0;
    for (int32_t p = 0; p < np; ++p) {
      for (int32_t i = b.prog_start[p]; i < b.prog_start[p + 1]; ++i) {
        uint32_t w = b.code[i];
        if (uses_const(w)) {
          const uint32_t arg = (w >> 8) + kofs;
          if (arg >= (1u << 24)) throw std::length_error("merged constant pool exceeds the 24-bit operand range");
          w = (w & 0xffu) | (arg << 8);
        }
        code.push_back(w);
      }
      prog_start.push_back((int32_t)code.size());
    }
    for (int32_t c = 0; c < b.n_cands; ++c) {
      cand_start.push_back(pbase + b.cand_start[c + 1]);
      handles.push_back(pv.dataset);
    }
  }
}

// ------------------------------------------------------------------ dispatcher
struct Dispatcher::Req {
  enum Kind { UPLOAD, RELEASE, FITNESS } kind;
  const double* X = nullptr;
  const double* y = nullptr;
  int64_t n = 0;
  int32_t d = 0;
  int32_t handle = -1;
  const rr_batch* b = nullptr;
  int32_t mode = 0, precision = 0;
  double* coefs = nullptr;
  rr_fit* fit = nullptr;
  double kernel_ms = 0;
  std::string error;
  bool done = false;
};

Dispatcher::Dispatcher(int device, int64_t slab_budget_mb, int gather_us) : gather_us_(gather_us > 0 ? gather_us : 0) {
  th_ = std::thread([this, device, slab_budget_mb] {
    std::string err;
    if (rr_create(device, &ctx_) != RR_OK) err = std::string("rr_create: ") + rr_last_error();
    else if (slab_budget_mb > 0 && rr_set_option(ctx_, "slab_budget_mb", slab_budget_mb) != RR_OK)
      err = std::string("rr_set_option: ") + rr_last_error();
    {
      std::lock_guard<std::mutex> g(m_);
      init_error_ = err;
      ready_ = true;
    }
    cv_.notify_all();
    if (err.empty()) loop();
  });
  std::unique_lock<std::mutex> lk(m_);
  cv_.wait(lk, [this] { return ready_; });
  if (!init_error_.empty()) {
    lk.unlock();
    th_.join();
    throw std::runtime_error(init_error_);
  }
}

Dispatcher::~Dispatcher() {
  {
    std::lock_guard<std::mutex> g(m_);
    stop_ = true;
  }
  cv_.notify_all();
  if (th_.joinable()) th_.join();
  if (ctx_) rr_destroy(ctx_);
}

Dispatcher::Counters Dispatcher::counters() {
  std::lock_guard<std::mutex> g(m_);
  return ctr_;
}

// A merged launch: the requests it carries and the arrays it needs until
// its results are handed out.
struct Dispatcher::Flight {
  std::vector<Req*> group;
  std::vector<int32_t> cs, ps, hs;
  std::vector<uint32_t> code;
  std::vector<double> consts, coefs;
  std::vector<rr_fit> fit;
  rr_batch mb{};
  int64_t ncand = 0;
};

// Merges the group's batches into one launch; false (with every request's
// error set) when they do not fit one.
bool Dispatcher::merge(Flight& f) {
  std::vector<BatchView> parts;
  for (Req* r : f.group) {
    parts.push_back({r->b, r->handle});
    f.ncand += r->b->n_cands;
  }
  try {
    merge_batches(parts, f.cs, f.ps, f.code, f.consts, f.hs);
  } catch (const std::exception& e) {
    for (Req* r : f.group) r->error = e.what();
    return false;
  }
  f.mb.n_cands = (int32_t)f.ncand;
  f.mb.cand_start = f.cs.data();
  f.mb.prog_start = f.ps.data();
  f.mb.code = f.code.data();
  f.mb.consts = f.consts.data();
  f.mb.n_consts = (int32_t)f.consts.size();
  f.coefs.assign((size_t)f.cs.back() + (size_t)f.ncand + 1, 0.0);
  f.fit.resize((size_t)std::max<int64_t>(f.ncand, 1));
  return true;
}

// Hands a finished launch's results (or its error) to its requests and wakes them.
void Dispatcher::finish(Flight& f, int rc, double ms) {
  if (rc != RR_OK) {
    const std::string e = rr_last_error();
    for (Req* r : f.group) if (r->error.empty()) r->error = e;
  } else {
    int64_t cbase = 0;
    for (Req* r : f.group) {
      const rr_batch& b = *r->b;
      for (int32_t c = 0; c < b.n_cands; ++c) {
        r->fit[c] = f.fit[(size_t)(cbase + c)];
        if (r->coefs) {
          const int64_t src = (int64_t)f.cs[(size_t)(cbase + c)] + cbase + c;
          const int64_t len = (int64_t)(b.cand_start[c + 1] - b.cand_start[c]) + 1;
          std::copy(f.coefs.begin() + src, f.coefs.begin() + src + len, r->coefs + b.cand_start[c] + c);
        }
      }
      r->kernel_ms = ms;
      cbase += b.n_cands;
    }
  }
  {
    std::lock_guard<std::mutex> g(m_);
    if (rc == RR_OK && f.ncand > 0) {
      ctr_.launches++;
      ctr_.requests += (int64_t)f.group.size();
      ctr_.candidates += f.ncand;
      ctr_.max_requests_per_launch = std::max<int64_t>(ctr_.max_requests_per_launch, (int64_t)f.group.size());
      ctr_.kernel_ms += ms;
    }
    for (Req* r : f.group) r->done = true;
  }
  cv_.notify_all();
}

// Waits for the oldest launch in flight and hands out its results.
void Dispatcher::retire_oldest() {
  std::unique_ptr<Flight> f = std::move(flights_.front());
  flights_.pop_front();
  const int rc = rr_fitness_wait(ctx_);
  double ms = 0;
  rr_last_kernel_ms(ctx_, &ms);
  finish(*f, rc, ms);
}

// The plain FP64 launches are pipelined: a merged launch is submitted
// (checked, planned and queued on the context's stream) while the previous
// one runs, and only then is the previous one waited for, so the host merge
// and check of one launch overlap the kernels of the other.  Other forms run
// one call at a time.
void Dispatcher::loop() {
  for (;;) {
    std::vector<Req*> batch;
    double gathered = 0;
    {
      std::unique_lock<std::mutex> lk(m_);
      if (flights_.empty()) cv_.wait(lk, [this] { return stop_ || !q_.empty(); });
      if (q_.empty() && stop_ && flights_.empty()) return;
      if (!q_.empty()) {
        // Continuous batching: a run whose sweep is not pending yet is on its
        // host work between sweeps; when the pending fitness requests are
        // fewer than the runs searching (those with a launch in flight
        // aside), give the others up to gather_us to join this launch
        // (uploads and releases never wait).
        auto fitness_pending = [this] {
          int f = 0;
          for (const Req* r : q_) {
            if (r->kind != Req::FITNESS) return -1;   // an upload or release: go now
            ++f;
          }
          return f;
        };
        int inflight = 0;
        for (const std::unique_ptr<Flight>& f : flights_) inflight += (int)f->group.size();
        const int searching = live_ - inflight;
        const int f0 = fitness_pending();
        if (gather_us_ > 0 && f0 > 0 && f0 < searching && !stop_) {
          const auto t0 = std::chrono::steady_clock::now();
          cv_.wait_until(lk, t0 + std::chrono::microseconds(gather_us_), [&] {
            const int f = fitness_pending();
            return stop_ || f < 0 || f >= searching;
          });
          gathered = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        batch.assign(q_.begin(), q_.end());   // everything pending now goes into this round
        q_.clear();
      }
    }
    const auto t0 = std::chrono::steady_clock::now();
    if (batch.empty()) {   // nothing new while a launch runs: retire it
      retire_oldest();
      continue;
    }
    int dlive = 0;
    std::vector<Req*> admin;
    for (Req* r : batch) if (r->kind != Req::FITNESS) admin.push_back(r);
    if (!admin.empty())
      while (!flights_.empty()) retire_oldest();   // uploads and releases need the context idle
    for (Req* r : admin) {
      if (r->kind == Req::UPLOAD) {
        if (rr_dataset_upload(ctx_, r->X, r->y, r->n, r->d, &r->handle) != RR_OK) r->error = rr_last_error();
        else ++dlive;
      } else {
        rr_dataset_free(ctx_, r->handle);
        --dlive;
      }
    }
    // fitness requests grouped by (mode, precision), in arrival order
    std::vector<Req*> fits;
    for (Req* r : batch) if (r->kind == Req::FITNESS) fits.push_back(r);
    while (!fits.empty()) {
      const int32_t mode = fits[0]->mode, prec = fits[0]->precision;
      std::vector<Req*> group, rest;
      for (Req* r : fits) (r->mode == mode && r->precision == prec ? group : rest).push_back(r);
      fits.swap(rest);
      std::unique_ptr<Flight> f(new Flight);
      f->group = std::move(group);
      if (!merge(*f)) {   // every request of the group already carries the error
        finish(*f, RR_E_LIMIT, 0.0);
        continue;
      }
      if (prec != RR_FP64) {
        while (!flights_.empty()) retire_oldest();
        const int rc = rr_fitness_multi(ctx_, f->hs.data(), &f->mb, mode, prec, f->coefs.data(), f->fit.data());
        double ms = 0;
        rr_last_kernel_ms(ctx_, &ms);
        finish(*f, rc, ms);
        continue;
      }
      if (flights_.size() == 2) retire_oldest();
      const int rc = rr_fitness_submit_multi(ctx_, f->hs.data(), &f->mb, mode, f->coefs.data(), f->fit.data());
      if (rc != RR_OK) {
        finish(*f, rc, 0.0);
        continue;
      }
      flights_.push_back(std::move(f));
      if (flights_.size() == 2) retire_oldest();   // the one that ran while this one was prepared
    }
    const double busy = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    {
      std::lock_guard<std::mutex> g(m_);
      ctr_.busy_s += busy;
      ctr_.gather_s += gathered;
      live_ += dlive;
      for (Req* r : admin) r->done = true;
    }
    cv_.notify_all();
  }
}

// Requests from the search threads: queue and block until the round that
// carries them has finished.
int32_t Dispatcher::upload(const double* X_fm, const double* y, int64_t n, int32_t d) {
  Req r;
  r.kind = Req::UPLOAD;
  r.X = X_fm; r.y = y; r.n = n; r.d = d;
  std::unique_lock<std::mutex> lk(m_);
  q_.push_back(&r);
  cv_.notify_all();
  cv_.wait(lk, [&] { return r.done; });
  if (!r.error.empty()) throw std::runtime_error("rr_dataset_upload: " + r.error);
  return r.handle;
}

void Dispatcher::release(int32_t handle) {
  Req r;
  r.kind = Req::RELEASE;
  r.handle = handle;
  std::unique_lock<std::mutex> lk(m_);
  q_.push_back(&r);
  cv_.notify_all();
  cv_.wait(lk, [&] { return r.done; });
}

void Dispatcher::fitness(int32_t handle, const rr_batch& b, int32_t mode, int32_t precision, double* coefs,
                         rr_fit* fit, double* kernel_ms) {
  Req r;
  r.kind = Req::FITNESS;
  r.handle = handle; r.b = &b; r.mode = mode; r.precision = precision; r.coefs = coefs; r.fit = fit;
  std::unique_lock<std::mutex> lk(m_);
  q_.push_back(&r);
  cv_.notify_all();
  cv_.wait(lk, [&] { return r.done; });
  if (!r.error.empty()) throw std::runtime_error("rr_fitness_multi: " + r.error);
  if (kernel_ms) *kernel_ms = r.kernel_ms;
}

// ------------------------------------------------------------------ server
struct Server::Run {
  int id = 0;
  RunSpec spec;
  RunResult res;
  bool done = false;
  std::thread th;
};

Server::Server(int device, int64_t hbm_budget_mb, int max_active, int host_threads_per_run, int64_t slab_budget_mb,
               int gather_us)
    : disp_(new Dispatcher(device, slab_budget_mb, gather_us)),
      budget_((hbm_budget_mb > 0 ? hbm_budget_mb : 200 * 1024) << 20),
      max_active_(max_active > 0 ? max_active : 1),
      threads_per_run_(host_threads_per_run > 0 ? host_threads_per_run : 1),
      t0_(std::chrono::steady_clock::now()) {}

Server::~Server() {
  std::vector<int> ids;
  {
    std::lock_guard<std::mutex> g(m_);
    for (auto& kv : runs_) ids.push_back(kv.first);
  }
  for (int id : ids) wait(id);
}

double Server::now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count(); }

int64_t Server::footprint(int64_t n_samples, int d) {
  const int64_t ldx = (n_samples + 4095) / 4096 * 4096;
  return ldx * std::max(d, 1) * (8 + 4) + ldx * 16;   // X (FP64 + FP32), y, the column of ones
}

int64_t Server::resident_bytes() {
  std::lock_guard<std::mutex> g(m_);
  return used_;
}

int Server::submit(RunSpec spec) {
  if ((int64_t)spec.X.size() != (int64_t)spec.n * spec.d || (int64_t)spec.y.size() != spec.n)
    throw std::invalid_argument("X must be n x d and y of length n");
  std::lock_guard<std::mutex> g(m_);
  const int id = next_id_++;
  std::unique_ptr<Run> r(new Run);
  r->id = id;
  const int64_t samples = std::max<int64_t>((int64_t)(spec.sample_size * (double)spec.n), 1);
  r->res.footprint = footprint(samples, spec.d);
  r->res.id = id;
  r->res.submitted_s = now();
  r->spec = std::move(spec);
  runs_[id] = std::move(r);
  pending_.push_back(id);
  admit_locked();
  return id;
}

void Server::admit_locked() {
  while (!pending_.empty()) {
    Run* r = runs_[pending_.front()].get();
    const bool fits = used_ + r->res.footprint <= budget_ || active_ == 0;
    if (active_ >= max_active_ || !fits) break;
    pending_.pop_front();
    ++active_;
    used_ += r->res.footprint;
    r->res.admitted_s = now();
    r->th = std::thread(&Server::run_body, this, r);
  }
}

void Server::run_body(Run* r) {
  const RunSpec& s = r->spec;
  RunResult out = r->res;
  try {
    RilsRols rr(s.classification, s.max_fit_calls, s.max_seconds, s.complexity_penalty, s.max_complexity,
                s.sample_size, false, s.random_state);
    rr.set_backend(disp_.get());
    rr.set_host_threads(threads_per_run_);
    rr.set_precision(s.precision);
    rr.fit(s.X.data(), (int64_t)s.X.size(), s.y.data(), (int64_t)s.y.size(), s.n, s.d);
    out.ok = true;
    out.model = rr.get_model_string();
    out.prefix = rr.final_prefix();
    out.fit_calls = rr.get_fit_calls();
    out.moves = rr.moves();
    out.sweep_ms = rr.sweep_ms();
    out.stats = rr.stats();
  } catch (const std::exception& e) {
    out.ok = false;
    out.error = e.what();
  }
  std::lock_guard<std::mutex> g(m_);
  out.finished_s = now();
  r->res = std::move(out);
  r->done = true;
  --active_;
  used_ -= r->res.footprint;
  admit_locked();
  cv_.notify_all();
}

bool Server::done(int id) {
  std::lock_guard<std::mutex> g(m_);
  auto it = runs_.find(id);
  if (it == runs_.end()) throw std::out_of_range("unknown run id");
  return it->second->done;
}

RunResult Server::wait(int id) {
  std::unique_lock<std::mutex> lk(m_);
  auto it = runs_.find(id);
  if (it == runs_.end()) throw std::out_of_range("unknown run id");
  Run* r = it->second.get();
  cv_.wait(lk, [&] { return r->done; });
  RunResult res = r->res;
  std::thread th = std::move(r->th);
  runs_.erase(it);
  lk.unlock();
  if (th.joinable()) th.join();
  return res;
}

}  // namespace rrh
