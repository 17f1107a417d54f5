// Host thread pools: a guard that joins every started thread of a pool when it goes out of
// scope, also while an exception (a thread that could not be started) unwinds, since a
// joinable std::thread must not be destroyed.
#pragma once

#include <thread>
#include <vector>

namespace sk {

struct JoinAll {
  std::vector<std::thread>& pool;
  ~JoinAll() {
    for (std::thread& th : pool)
      if (th.joinable()) th.join();
  }
};

}  // namespace sk
