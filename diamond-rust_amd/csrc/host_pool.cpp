#include "host_pool.hpp"

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

namespace dmr {

namespace {

class HostPool {
 public:
  HostPool() {
    const unsigned hw = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    for (unsigned t = 0; t + 1 < hw; ++t) workers_.emplace_back([this] { work(); });
  }
  ~HostPool() {
    {
      std::lock_guard<std::mutex> g(m_);
      stop_ = true;
    }
    wake_.notify_all();
    for (auto& w : workers_) w.join();
  }
  unsigned threads() const { return (unsigned)workers_.size() + 1; }

  // f(part) for part = 0..parts-1; the caller takes parts too.
  void run(unsigned parts, const std::function<void(unsigned)>& f) {
    std::lock_guard<std::mutex> one(run_m_);
    {
      std::lock_guard<std::mutex> g(m_);
      f_ = &f;
      parts_ = parts;
      next_.store(0);
      busy_ = (unsigned)workers_.size();
      ++generation_;
    }
    wake_.notify_all();
    take();
    std::unique_lock<std::mutex> g(m_);
    done_.wait(g, [this] { return busy_ == 0; });
    f_ = nullptr;
  }

 private:
  void take() {
    for (unsigned p; (p = next_.fetch_add(1)) < parts_;) (*f_)(p);
  }
  void work() {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> g(m_);
        wake_.wait(g, [&] { return stop_ || generation_ != seen; });
        if (stop_) return;
        seen = generation_;
      }
      take();
      std::lock_guard<std::mutex> g(m_);
      if (--busy_ == 0) done_.notify_one();
    }
  }

  std::vector<std::thread> workers_;
  std::mutex m_, run_m_;
  std::condition_variable wake_, done_;
  const std::function<void(unsigned)>* f_ = nullptr;
  unsigned parts_ = 0, busy_ = 0;
  std::atomic<unsigned> next_{0};
  uint64_t generation_ = 0;
  bool stop_ = false;
};

HostPool& pool() {
  static HostPool p;
  return p;
}

}  // namespace

unsigned host_threads() { return pool().threads(); }

void host_parallel_for(uint64_t n, uint64_t min_per_part,
                       const std::function<void(uint64_t, uint64_t)>& f) {
  if (n == 0) return;
  const uint64_t want = (n + std::max<uint64_t>(1, min_per_part) - 1) / std::max<uint64_t>(1, min_per_part);
  if (want <= 1) {
    f(0, n);
    return;
  }
  HostPool& p = pool();
  const unsigned parts = (unsigned)std::min<uint64_t>(want, p.threads());
  if (parts <= 1) {
    f(0, n);
    return;
  }
  p.run(parts, [&](unsigned k) { f(n * k / parts, n * (k + 1) / parts); });
}

}  // namespace dmr
