// Host-side helpers shared by the library's entry points.
#pragma once
#include <algorithm>
#include <chrono>
#include <cstddef>
#include <thread>
#include <vector>

namespace dnab {

inline double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// fn(b, e) over [0, n) split into up to 16 host threads (at least 256 items
// each); one call on the caller's thread when n is small.
template <class F>
void parallel_for(size_t n, F fn) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t nt = std::min<size_t>(std::min<unsigned>(hw, 16u), std::max<size_t>(1, n / 256));
  if (nt <= 1) {
    fn(size_t{0}, n);
    return;
  }
  std::vector<std::thread> th;
  const size_t per = (n + nt - 1) / nt;
  for (size_t k = 0; k < nt; ++k) {
    const size_t b = std::min(n, k * per), e = std::min(n, b + per);
    if (b < e) th.emplace_back(fn, b, e);
  }
  for (auto& t : th) t.join();
}

}  // namespace dnab
