// Checks of the host thread pool (anukulan_amd/csrc/host_pool.cpp), built and
// run by tests/test_host_pool.py: every index runs once, an exception in a
// loop body reaches the caller, a loop started inside a loop body runs on
// its caller, and loops from two threads at once both complete.
#include <atomic>
#include <cstdio>
#include <stdexcept>
#include <thread>
#include <vector>

#include "host_pool.hpp"

using anuk::HostPool;

int main() {
  HostPool& pool = HostPool::shared();
  std::printf("threads %zu\n", pool.threads());
  int failures = 0;
  auto expect = [&](bool ok, const char* what) {
    if (!ok) {
      std::printf("FAIL %s\n", what);
      ++failures;
    }
  };
  for (std::size_t count : {0u, 1u, 2u, 15u, 16u, 17u, 1000u, 100000u}) {
    std::vector<std::atomic<int>> hits(count);
    for (auto& h : hits) h.store(0);
    pool.run(count, [&](std::size_t i) { hits[i].fetch_add(1); });
    bool once = true;
    for (auto& h : hits) once = once && h.load() == 1;
    expect(once, "every index once");
  }
  bool caught = false;
  try {
    pool.run(1000, [&](std::size_t i) {
      if (i == 517) throw std::runtime_error("index 517");
    });
  } catch (const std::runtime_error& e) {
    caught = std::string(e.what()) == "index 517";
  }
  expect(caught, "exception reaches the caller");
  std::atomic<long> total{0};
  pool.run(64, [&](std::size_t i) {
    pool.run(64, [&](std::size_t j) { total.fetch_add(static_cast<long>(i * 64 + j)); });
  });
  expect(total.load() == 4095L * 4096L / 2L, "nested loops");
  std::atomic<long> a{0}, b{0};
  std::thread other([&] {
    for (int r = 0; r < 50; ++r) pool.run(5000, [&](std::size_t i) { a.fetch_add(static_cast<long>(i)); });
  });
  for (int r = 0; r < 50; ++r) pool.run(5000, [&](std::size_t i) { b.fetch_add(static_cast<long>(i)); });
  other.join();
  expect(a.load() == 50L * 4999L * 5000L / 2L && b.load() == a.load(), "two threads at once");
  std::printf(failures == 0 ? "ok\n" : "failed\n");
  return failures == 0 ? 0 : 1;
}
