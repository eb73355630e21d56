// A process-wide pool of host worker threads for short parallel regions
// (batch packing, staging copies).  Creating 16 threads per region cost more
// than the work of a 20k-pair batch; the pool's threads are created once and
// woken per region.
#pragma once
#include <algorithm>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace hostpool {

class Pool {
 public:
  static constexpr int kMaxThreads = 16;

  // fn(t) for t in [0, nth) on nth threads (the caller runs t = 0), returning
  // when all are done.  One region at a time: a caller that finds the pool
  // busy runs its region on threads of its own.
  void run(int nth, const std::function<void(int)>& fn) { region(nth, fn, false); }

  // For regions that share their work out through a counter of their own:
  // fn(t) runs on the caller (t = 0) and on each worker t < nth that wakes
  // before the caller's fn(0) has returned, i.e. while work may be left.
  // The call returns once those have finished; a worker that wakes later
  // skips the region, so a slow wake-up never holds the caller.
  void run_claimed(int nth, const std::function<void(int)>& fn) { region(nth, fn, true); }

  ~Pool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_work_.notify_all();
    for (auto& t : workers_) t.join();
  }

 private:
  void region(int nth, const std::function<void(int)>& fn, bool claimed) {
    nth = std::max(1, std::min(nth, kMaxThreads));
    if (nth == 1) {
      fn(0);
      return;
    }
    std::unique_lock<std::mutex> region(region_mu_, std::try_to_lock);
    if (!region.owns_lock()) {
      std::vector<std::thread> th;
      for (int t = 1; t < nth; t++) th.emplace_back(fn, t);
      fn(0);
      for (auto& t : th) t.join();
      return;
    }
    {
      std::unique_lock<std::mutex> lk(mu_);
      while ((int)workers_.size() < nth - 1) {
        const int id = (int)workers_.size() + 1;
        workers_.emplace_back([this, id] { worker(id); });
      }
      fn_ = &fn;
      active_ = nth;
      open_ = true;
      running_ = 0;
      pending_ = claimed ? 0 : nth - 1;
      generation_++;
    }
    cv_work_.notify_all();
    fn(0);
    std::unique_lock<std::mutex> lk(mu_);
    if (claimed) open_ = false;  // workers not yet awake skip the region
    cv_done_.wait(lk, [&] { return pending_ == 0 && running_ == 0; });
    open_ = false;
    fn_ = nullptr;
  }

  void worker(int id) {
    uint64_t seen = 0;
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      cv_work_.wait(lk, [&] { return stop_ || generation_ != seen; });
      if (stop_) return;
      seen = generation_;
      if (id >= active_ || !open_) continue;  // not part of this region, or a claimed region already closed
      const bool counted = pending_ > 0;       // a region where every index runs
      running_++;
      const std::function<void(int)>* fn = fn_;
      lk.unlock();
      (*fn)(id);
      lk.lock();
      running_--;
      if (counted) pending_--;
      if (pending_ == 0 && running_ == 0) cv_done_.notify_one();
    }
  }

  std::mutex region_mu_;  // one region at a time
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::vector<std::thread> workers_;
  const std::function<void(int)>* fn_ = nullptr;
  int active_ = 0, pending_ = 0, running_ = 0;
  bool open_ = false;
  uint64_t generation_ = 0;
  bool stop_ = false;
};

inline Pool& pool() {
  static Pool p;
  return p;
}

}  // namespace hostpool
