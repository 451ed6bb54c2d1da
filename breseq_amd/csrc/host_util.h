// host_util.h -- shared host-side helpers: error type, status filling, number formatting that
// reproduces breseq's iostream output, and the host trims used when synthesizing XL/XR tags.
#pragma once

#include <algorithm>
#include <cstdint>
#include <functional>
#include <future>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bsq.h"

namespace bsq {

struct BsqError : std::runtime_error {
  int code;
  BsqError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_status(bsq_status* st, int code, const std::string& msg);

// breseq to_string(double, precision, scientific) (common.h:845-866): "NA" for NaN, fixed or
// scientific with `precision` digits after the point.
std::string fmt_double(double v, unsigned precision, bool scientific);
// fmt_double into buf (NUL-terminated); returns the length, or -1 if buf is too small.
int fmt_double_buf(double v, unsigned precision, bool scientific, char* buf, size_t len);
// iostream default formatting of a double (%g, 6 significant digits).
std::string fmt_g(double v, int precision = 6);
// The double that reading fmt_g(v) back with strtod gives, without the text: exact 128-bit
// integer arithmetic finds the six significant digits w and the scale k that %g prints (rounded half
// to even on v's exact binary value), and w / 10^k or w * 10^-k is then one correctly rounded IEEE
// operation, since w < 2^53 and 10^|k| <= 10^22 are exact doubles (Clinger's fast path of
// decimal-to-binary conversion).  Values it does not cover (0, non-finite, |v| outside
// [1e-9, 1e15)) go through the text.
double g6_round_trip(double v);

// Host threads for parallel host work: the hardware threads, at most 16.
inline int host_threads() {
  const unsigned hc = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(hc == 0 ? 4u : hc, 16u));
}

// Runs fn(t, begin, end) over [0, n) split into contiguous ranges, one per thread; an exception
// thrown in a thread is rethrown here as a BsqError.
template <class F>
void parallel_ranges(int64_t n, int threads, F fn) {
  threads = (int)std::max<int64_t>(1, std::min<int64_t>(threads, n));
  std::vector<std::thread> pool;
  std::vector<std::string> errors(threads);
  std::vector<int> codes(threads, 0);
  for (int t = 0; t < threads; t++) {
    const int64_t b = n * t / threads, e = n * (t + 1) / threads;
    pool.emplace_back([&, t, b, e] {
      try {
        fn(t, b, e);
      } catch (const BsqError& ex) {
        codes[t] = ex.code;
        errors[t] = ex.what();
      } catch (const std::exception& ex) {
        codes[t] = BSQ_ERR_FORMAT;
        errors[t] = ex.what();
      }
    });
  }
  for (auto& th : pool) th.join();
  for (int t = 0; t < threads; t++)
    if (codes[t]) throw BsqError(codes[t], errors[t]);
}

// Runs fn(i) for every task i in [0, n_tasks) on a pool of host threads that lives for the
// process (host_threads() - 1 workers plus the calling thread), and returns when every task has
// ended.  Several callers may submit at once: each takes part in its own job, so a busy pool
// never blocks one.  An exception in a task is rethrown here as a BsqError.
void pool_run(int n_tasks, const std::function<void(int)>& fn);

// Runs task on a host thread that lives for the process and returns at once: a thread is started
// only when every existing one is busy, so host work beside the device passes pays no thread start
// per call.
void async_submit(std::function<void()> task);

// fn() on such a thread; the future holds its result or its exception.  Unlike std::async's, the
// future does not wait in its destructor: callers whose fn refers to their locals wait for it
// (WaitOnExit) before those go.
template <class F>
auto run_async(F fn) -> std::future<decltype(fn())> {
  using R = decltype(fn());
  auto task = std::make_shared<std::packaged_task<R()>>(std::move(fn));
  std::future<R> f = task->get_future();
  async_submit([task] { (*task)(); });
  return f;
}

// Waits for a future (or shared future) when the scope ends, also when it ends by an exception.
template <class Fut>
struct WaitOnExit {
  Fut& f;
  ~WaitOnExit() {
    if (f.valid()) f.wait();
  }
};

// SequenceTrims (calculate_trims.cpp:67-138): out[0..len) left, out[len..2len) right.
void host_sequence_trims(const char* seq, int64_t len, uint8_t* out);
// edge_trims_for_sequence (calculate_trims.cpp:183-230) of a read's aligned bases.
void host_edge_trims(const std::string& seq, uint32_t& left, uint32_t& right);

}  // namespace bsq
