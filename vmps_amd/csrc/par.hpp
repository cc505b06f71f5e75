// Host-side parallel loop for the plan builders' bulk copies (environment
// gathers, split-BF16 operand conversion).
#pragma once
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <thread>
#include <vector>

namespace vmps {

// Host threads used for bulk copies: the hardware concurrency, at most 16
// (the share of a GPU's host cores this library assumes).
inline unsigned host_threads() {
  const unsigned hc = std::thread::hardware_concurrency();
  return std::max(1u, std::min(16u, hc ? hc : 1u));
}

// fn(i) for i in [0, n), items handed out one at a time to host_threads()
// threads; serial when `work` (elements touched) is below 4M.  fn must not
// throw.
template <class F>
void parallel_for(int64_t n, int64_t work, F fn) {
  const unsigned nt = host_threads();
  if (n < 2 || nt < 2 || work < (int64_t(1) << 22)) {
    for (int64_t i = 0; i < n; ++i) fn(i);
    return;
  }
  std::atomic<int64_t> next{0};
  auto worker = [&] {
    for (int64_t i; (i = next.fetch_add(1, std::memory_order_relaxed)) < n;) fn(i);
  };
  std::vector<std::thread> pool;
  const unsigned k = (unsigned)std::min<int64_t>(nt, n);
  for (unsigned t = 1; t < k; ++t) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
}

}  // namespace vmps
