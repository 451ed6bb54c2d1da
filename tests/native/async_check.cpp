// run_async / async_submit (host_util.cpp): results and exceptions reach the futures, tasks that
// wait on each other all run (a thread is started when none is idle), and WaitOnExit waits.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <stdexcept>
#include <thread>
#include <vector>

#include "../../breseq_amd/csrc/host_util.h"

int main() {
  using namespace bsq;
  int bad = 0;
  // results, many at once
  std::vector<std::future<int>> fs;
  for (int i = 0; i < 200; i++) fs.push_back(run_async([i] { return i * i; }));
  for (int i = 0; i < 200; i++)
    if (fs[i].get() != i * i) bad++;
  // an exception reaches its future
  auto fe = run_async([]() -> int { throw std::runtime_error("boom"); });
  try {
    fe.get();
    bad++;
  } catch (const std::runtime_error&) {
  }
  // 16 tasks that all wait for each other: each needs a thread of its own
  std::atomic<int> arrived{0};
  std::vector<std::future<void>> gs;
  for (int i = 0; i < 16; i++)
    gs.push_back(run_async([&] {
      arrived++;
      const auto t0 = std::chrono::steady_clock::now();
      while (arrived.load() < 16 && std::chrono::steady_clock::now() - t0 < std::chrono::seconds(20))
        std::this_thread::yield();
    }));
  for (auto& g : gs) g.get();
  if (arrived.load() != 16) bad++;
  // WaitOnExit: the task has finished when the scope ends
  std::atomic<int> done{0};
  {
    std::shared_future<void> f = run_async([&] {
      std::this_thread::sleep_for(std::chrono::milliseconds(50));
      done = 1;
    }).share();
    WaitOnExit<std::shared_future<void>> w{f};
  }
  if (done.load() != 1) bad++;
  printf("%d failures\n", bad);
  return bad == 0 ? 0 : 1;
}
