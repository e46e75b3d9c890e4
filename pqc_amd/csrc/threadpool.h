// A fixed set of host threads for the library's copy loops (host-buffer staging,
// the serving loop's batch packing and result delivery), started once so that a
// copy never pays for thread creation.
#pragma once
#include <algorithm>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace pqc {

// run(fn) calls fn(t, T) on T threads (the caller is thread 0) and returns when
// every call has returned.  Calls from several threads are serialised.
class CopyPool {
 public:
  explicit CopyPool(int threads) : n_(threads < 1 ? 1 : threads) {
    for (int t = 1; t < n_; ++t) th_.emplace_back([this, t] { worker(t); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  int size() const { return n_; }
  void run(const std::function<void(int, int)>& fn) {
    std::lock_guard<std::mutex> one(run_mu_);
    {
      std::lock_guard<std::mutex> lk(mu_);
      job_ = &fn;
      pending_ = n_ - 1;
      ++gen_;
    }
    cv_.notify_all();
    fn(0, n_);
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [this] { return pending_ == 0; });
    job_ = nullptr;
  }
  // fn(begin, end) over [0, n) in contiguous chunks of at least `per` items, on
  // as many of the pool's threads as that gives (one chunk: the caller alone)
  template <typename F>
  void parallel_for(size_t n, size_t per, F fn) {
    const int want = (int)std::min<size_t>((size_t)n_, (n + per - 1) / std::max<size_t>(per, 1));
    if (want <= 1) {
      if (n) fn((size_t)0, n);
      return;
    }
    const size_t chunk = (n + want - 1) / want;
    run([&](int t, int) {
      const size_t b = std::min(n, (size_t)t * chunk), e = std::min(n, b + chunk);
      if (t < want && b < e) fn(b, e);
    });
  }

 private:
  void worker(int t) {
    uint64_t seen = 0;
    for (;;) {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      const std::function<void(int, int)>* f = job_;
      lk.unlock();
      (*f)(t, n_);
      lk.lock();
      if (--pending_ == 0) done_.notify_one();
    }
  }
  int n_;
  std::vector<std::thread> th_;
  std::mutex mu_, run_mu_;
  std::condition_variable cv_, done_;
  const std::function<void(int, int)>* job_ = nullptr;
  uint64_t gen_ = 0;
  int pending_ = 0;
  bool stop_ = false;
};

}  // namespace pqc
