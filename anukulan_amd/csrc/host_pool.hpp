// A fixed set of host threads for the solvers' per-lane host work (staging
// the lanes' vectors, replaying scaling sweeps, residuals of results):
// started on first use and kept, so a parallel loop costs a wake-up rather
// than thread creation.
#pragma once

#include <cstddef>
#include <functional>

namespace anuk {

class HostPool {
 public:
  // The process-wide pool: min(16, hardware threads) threads in all, the
  // caller counted.
  static HostPool& shared();
  // Runs f(i) for every i in [0, count), spread over the pool's threads and
  // the caller, and returns when all are done; the first exception a call
  // throws is rethrown here. A loop started while the pool is busy (from
  // another thread, or from inside a loop body) runs on the caller alone.
  void run(std::size_t count, const std::function<void(std::size_t)>& f);
  std::size_t threads() const;

  HostPool(const HostPool&) = delete;
  HostPool& operator=(const HostPool&) = delete;

 private:
  explicit HostPool(std::size_t threads);
  struct State;
  State* s_;
};

}  // namespace anuk
