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
      stop_flag_.store(true);
    }
    wake_.notify_all();
    for (auto& w : workers_) w.join();
  }
  unsigned threads() const { return (unsigned)workers_.size() + 1; }

  // f(part) for part = 0..parts-1; the caller takes parts too. Parts are
  // claimed one at a time from a shared counter, and the call waits only for
  // the workers that joined this run, so a worker that wakes late (after
  // every part was claimed) delays nothing.
  void run(unsigned parts, const std::function<void(unsigned)>& f) {
    std::lock_guard<std::mutex> one(run_m_);
    {
      std::lock_guard<std::mutex> g(m_);
      f_ = &f;
      parts_ = parts;
      next_.store(0);
      active_ = true;
      generation_.fetch_add(1);
    }
    wake_.notify_all();
    for (unsigned p; (p = next_.fetch_add(1)) < parts;) f(p);
    std::unique_lock<std::mutex> g(m_);
    done_.wait(g, [this] { return inflight_ == 0; });
    active_ = false;
    f_ = nullptr;
  }

 private:
  void work() {
    uint64_t seen = 0;
    for (;;) {
      // A short spin first: runs often come in quick succession.
      for (int k = 0; k < kSpin && generation_.load() == seen && !stop_flag_.load(); ++k)
        std::this_thread::yield();
      const std::function<void(unsigned)>* f;
      unsigned parts;
      {
        std::unique_lock<std::mutex> g(m_);
        wake_.wait(g, [&] { return stop_ || generation_.load() != seen; });
        if (stop_) return;
        seen = generation_.load();
        if (!active_) continue;
        f = f_;
        parts = parts_;
        ++inflight_;
      }
      for (unsigned p; (p = next_.fetch_add(1)) < parts;) (*f)(p);
      std::lock_guard<std::mutex> g(m_);
      if (--inflight_ == 0) done_.notify_one();
    }
  }

  static constexpr int kSpin = 500;
  std::vector<std::thread> workers_;
  std::mutex m_, run_m_;
  std::condition_variable wake_, done_;
  const std::function<void(unsigned)>* f_ = nullptr;
  unsigned parts_ = 0, inflight_ = 0;
  std::atomic<unsigned> next_{0};
  std::atomic<uint64_t> generation_{0};
  bool active_ = false, stop_ = false;
  std::atomic<bool> stop_flag_{false};
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
  // Up to four parts per thread, claimed as threads come free.
  const unsigned parts = (unsigned)std::min<uint64_t>(want, 4ull * p.threads());
  if (parts <= 1) {
    f(0, n);
    return;
  }
  p.run(parts, [&](unsigned k) { f(n * k / parts, n * (k + 1) / parts); });
}

}  // namespace dmr
