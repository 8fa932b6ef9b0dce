// workers.h -- a small persistent pool of host threads for the per-candidate
// host work of a sweep (rewrite rounds, factor extraction, strings).  Items
// are independent and write only their own outputs, so results do not depend
// on the thread count or the schedule.
#pragma once
#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace rrh {

class Workers {
 public:
  explicit Workers(int threads) : n_(threads < 1 ? 1 : threads) {
    for (int t = 1; t < n_; ++t) th_.emplace_back([this, t] { loop(t); });
  }
  ~Workers() {
    {
      std::lock_guard<std::mutex> g(m_);
      stop_ = true;
      ++gen_;
    }
    cv_.notify_all();
    for (std::thread& t : th_) t.join();
  }
  Workers(const Workers&) = delete;
  Workers& operator=(const Workers&) = delete;

  int size() const { return n_; }

  // f(i, tid) for every i in [0, count), tid in [0, size()); returns when all
  // items are done.  Small batches run on the calling thread.  Threads take
  // chunk consecutive items at a time.
  void parallel_for(size_t count, const std::function<void(size_t, int)>& f, size_t min_items = 32,
                    size_t chunk = 4) {
    if (n_ == 1 || count < min_items) {
      for (size_t i = 0; i < count; ++i) f(i, 0);
      return;
    }
    {
      std::lock_guard<std::mutex> g(m_);
      fn_ = &f;
      count_ = count;
      chunk_ = chunk < 1 ? 1 : chunk;
      next_.store(0);
      busy_ = n_ - 1;
      ++gen_;
    }
    cv_.notify_all();
    work(0);
    std::unique_lock<std::mutex> lk(m_);
    done_cv_.wait(lk, [this] { return busy_ == 0; });
    fn_ = nullptr;
  }

 private:
  void work(int tid) {
    for (;;) {
      const size_t b = next_.fetch_add(chunk_);
      if (b >= count_) break;
      const size_t e = b + chunk_ < count_ ? b + chunk_ : count_;
      for (size_t i = b; i < e; ++i) (*fn_)(i, tid);
    }
  }
  void loop(int tid) {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
        if (stop_) return;
      }
      work(tid);
      {
        std::lock_guard<std::mutex> g(m_);
        if (--busy_ == 0) done_cv_.notify_one();
      }
    }
  }

  const int n_;
  std::vector<std::thread> th_;
  std::mutex m_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(size_t, int)>* fn_ = nullptr;
  size_t count_ = 0;
  size_t chunk_ = 4;
  std::atomic<size_t> next_{0};
  int busy_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

}  // namespace rrh
