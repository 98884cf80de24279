#include "host_pool.hpp"

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

namespace anuk {

struct HostPool::State {
  std::size_t threads = 1;
  std::vector<std::thread> workers;
  std::mutex busy;  // held by the loop in progress
  std::mutex m;
  std::condition_variable wake, done;
  std::size_t generation = 0;
  std::size_t running = 0;  // workers still inside the current loop
  const std::function<void(std::size_t)>* job = nullptr;
  std::size_t count = 0;
  std::atomic<std::size_t> next{0};
  std::exception_ptr error;

  // Takes indices of the current loop until none are left.
  void drain() {
    for (;;) {
      const std::size_t i = next.fetch_add(1, std::memory_order_relaxed);
      if (i >= count) return;
      try {
        (*job)(i);
      } catch (...) {
        std::lock_guard<std::mutex> g(m);
        if (!error) error = std::current_exception();
        next.store(count, std::memory_order_relaxed);  // stop handing out work
      }
    }
  }

  void worker() {
    std::size_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> g(m);
        wake.wait(g, [&] { return generation != seen; });
        seen = generation;
      }
      drain();
      std::lock_guard<std::mutex> g(m);
      if (--running == 0) done.notify_one();
    }
  }
};

HostPool::HostPool(std::size_t threads) : s_(new State) {
  s_->threads = std::max<std::size_t>(1, threads);
  for (std::size_t t = 1; t < s_->threads; ++t) s_->workers.emplace_back([this] { s_->worker(); });
}

HostPool& HostPool::shared() {
  // Never destroyed: the workers stay parked until the process ends.
  static HostPool* pool = new HostPool(
      std::min<std::size_t>(16, std::max<unsigned>(1u, std::thread::hardware_concurrency())));
  return *pool;
}

std::size_t HostPool::threads() const { return s_->threads; }

void HostPool::run(std::size_t count, const std::function<void(std::size_t)>& f) {
  std::unique_lock<std::mutex> busy(s_->busy, std::try_to_lock);
  if (count <= 1 || s_->workers.empty() || !busy.owns_lock()) {
    for (std::size_t i = 0; i < count; ++i) f(i);
    return;
  }
  {
    std::lock_guard<std::mutex> g(s_->m);
    s_->job = &f;
    s_->count = count;
    s_->next.store(0, std::memory_order_relaxed);
    s_->error = nullptr;
    s_->running = s_->workers.size();
    ++s_->generation;
  }
  s_->wake.notify_all();
  s_->drain();
  std::exception_ptr error;
  {
    std::unique_lock<std::mutex> g(s_->m);
    s_->done.wait(g, [&] { return s_->running == 0; });
    s_->job = nullptr;
    error = s_->error;
  }
  if (error) std::rethrow_exception(error);
}

}  // namespace anuk
