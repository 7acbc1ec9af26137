// Host-side helpers shared by the library's entry points.
#pragma once
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstddef>
#include <functional>
#include <memory>
#include <mutex>
#include <utility>
#include <thread>
#include <vector>

namespace dnab {

// Allocator whose value-construction leaves elements uninitialised (for
// host arrays that are filled right after sizing them).
template <class T>
struct NoInitAlloc : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() = default;
  template <class U>
  NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
  template <class U>
  void construct(U* p) noexcept {
    ::new (static_cast<void*>(p)) U;
  }
  template <class U, class... A>
  void construct(U* p, A&&... a) {
    ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
  }
};
template <class T>
using HostVec = std::vector<T, NoInitAlloc<T>>;

inline double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Worker threads kept for the life of the process (the pool is never
// destroyed, so no thread is joined at exit). One job at a time: a job is
// parts 0..n_parts-1 of a function; the caller runs part 0 and waits for the
// rest. A call made while another job runs, or from inside a job, runs all
// its parts on the calling thread.
class HostPool {
 public:
  static HostPool& get() {
    static HostPool* pool = new HostPool();
    return *pool;
  }
  size_t workers() const { return th_.size(); }
  // false when the pool is busy (the caller then runs the parts itself)
  bool run(size_t n_parts, const std::function<void(size_t)>& part) {
    if (in_job_ || n_parts <= 1 || n_parts > th_.size() + 1) return false;
    std::unique_lock<std::mutex> job(job_mu_, std::try_to_lock);
    if (!job.owns_lock()) return false;
    {
      std::lock_guard<std::mutex> lk(mu_);
      part_ = &part;
      n_parts_ = n_parts;
      next_ = 1;
      left_ = n_parts_ - 1;
      ++gen_;
    }
    cv_.notify_all();
    in_job_ = true;
    part(0);
    in_job_ = false;
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return left_ == 0; });
    part_ = nullptr;
    return true;
  }

 private:
  HostPool() {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned n = std::min(hw, 16u) - 1;
    for (unsigned k = 0; k < n; ++k) th_.emplace_back([this] { loop(); });
    for (auto& t : th_) t.detach();
  }
  void loop() {
    in_job_ = true;  // a job's part that calls parallel_for runs it serially
    uint64_t seen = 0;
    for (;;) {
      size_t k;
      const std::function<void(size_t)>* f;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen && next_ < n_parts_; });
        k = next_++;
        if (next_ >= n_parts_) seen = gen_;
        f = part_;
      }
      (*f)(k);
      std::lock_guard<std::mutex> lk(mu_);
      if (--left_ == 0) done_.notify_one();
    }
  }
  std::vector<std::thread> th_;
  std::mutex job_mu_, mu_;
  std::condition_variable cv_, done_;
  const std::function<void(size_t)>* part_ = nullptr;
  size_t n_parts_ = 0, next_ = 0, left_ = 0;
  uint64_t gen_ = 0;
  static thread_local bool in_job_;
};
inline thread_local bool HostPool::in_job_ = false;

// fn(b, e) over [0, n) split over the pool's threads and the caller (up to
// 16 parts, at least 256 items each); on the caller's thread alone when n is
// small or the pool is busy.
template <class F>
void parallel_for(size_t n, F fn) {
  HostPool& pool = HostPool::get();
  const size_t np = std::min<size_t>(pool.workers() + 1, std::max<size_t>(1, n / 256));
  if (np <= 1) {
    fn(size_t{0}, n);
    return;
  }
  const size_t per = (n + np - 1) / np;
  const std::function<void(size_t)> part = [&](size_t k) {
    const size_t b = std::min(n, k * per), e = std::min(n, b + per);
    if (b < e) fn(b, e);
  };
  if (!pool.run(np, part))
    for (size_t k = 0; k < np; ++k) part(k);
}

}  // namespace dnab
