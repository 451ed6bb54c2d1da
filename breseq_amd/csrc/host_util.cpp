// host_util.cpp -- see host_util.h.
#include "host_util.h"

#include <charconv>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>

namespace bsq {

namespace {

// One submitted job: tasks are claimed through `next`; the submitter waits for `finished`.
struct PoolJob {
  int n = 0;
  const std::function<void(int)>* fn = nullptr;
  std::atomic<int> next{0};
  std::atomic<int> finished{0};
  std::mutex m;
  std::condition_variable cv;
  int err_code = 0;
  std::string err_msg;

  // Claims and runs tasks until none is left; returns the number run.
  int drain() {
    int ran = 0;
    for (int i = next++; i < n; i = next++) {
      try {
        (*fn)(i);
      } catch (const BsqError& e) {
        std::lock_guard<std::mutex> g(m);
        if (!err_code) {
          err_code = e.code;
          err_msg = e.what();
        }
      } catch (const std::exception& e) {
        std::lock_guard<std::mutex> g(m);
        if (!err_code) {
          err_code = BSQ_ERR_FORMAT;
          err_msg = e.what();
        }
      }
      ran++;
    }
    return ran;
  }
  void add_finished(int k) {
    if (k == 0) return;
    if (finished.fetch_add(k) + k == n) {
      std::lock_guard<std::mutex> g(m);
      cv.notify_all();
    }
  }
};

class Pool {
 public:
  static Pool& get() {
    static Pool* p = new Pool();   // never destroyed: workers may outlive static destructors
    return *p;
  }
  void run(int n_tasks, const std::function<void(int)>& fn) {
    if (n_tasks <= 0) return;
    auto job = std::make_shared<PoolJob>();
    job->n = n_tasks;
    job->fn = &fn;
    if (!workers_.empty() && n_tasks > 1) {
      std::lock_guard<std::mutex> g(m_);
      jobs_.push_back(job);
      cv_.notify_all();
    }
    job->add_finished(job->drain());
    {
      std::unique_lock<std::mutex> g(job->m);
      job->cv.wait(g, [&] { return job->finished.load() == job->n; });
    }
    if (job->err_code) throw BsqError(job->err_code, job->err_msg);
  }

 private:
  Pool() {
    const int n = host_threads() - 1;
    for (int i = 0; i < n; i++) workers_.emplace_back([this] { loop(); });
    for (auto& t : workers_) t.detach();
  }
  void loop() {
    for (;;) {
      std::shared_ptr<PoolJob> job;
      {
        std::unique_lock<std::mutex> g(m_);
        cv_.wait(g, [&] {
          while (!jobs_.empty() && jobs_.front()->next.load() >= jobs_.front()->n) jobs_.pop_front();
          return !jobs_.empty();
        });
        job = jobs_.front();
      }
      job->add_finished(job->drain());
    }
  }
  std::mutex m_;
  std::condition_variable cv_;
  std::deque<std::shared_ptr<PoolJob>> jobs_;
  std::vector<std::thread> workers_;
};

}  // namespace

void pool_run(int n_tasks, const std::function<void(int)>& fn) { Pool::get().run(n_tasks, fn); }

namespace {

// Threads for async_submit: a task goes to an idle thread, or to a new one when none is idle.
class AsyncWorkers {
 public:
  static AsyncWorkers& get() {
    static AsyncWorkers* w = new AsyncWorkers();   // never destroyed, like the pool
    return *w;
  }
  void submit(std::function<void()> task) {
    bool start = false;
    {
      std::lock_guard<std::mutex> g(m_);
      tasks_.push_back(std::move(task));
      if (idle_ < (int)tasks_.size()) {
        start = true;
        threads_++;
      }
    }
    if (start) std::thread([this] { loop(); }).detach();
    else cv_.notify_one();
  }

 private:
  void loop() {
    for (;;) {
      std::function<void()> task;
      {
        std::unique_lock<std::mutex> g(m_);
        idle_++;
        cv_.wait(g, [&] { return !tasks_.empty(); });
        idle_--;
        task = std::move(tasks_.front());
        tasks_.pop_front();
      }
      task();
    }
  }
  std::mutex m_;
  std::condition_variable cv_;
  std::deque<std::function<void()>> tasks_;
  int idle_ = 0, threads_ = 0;
};

}  // namespace

void async_submit(std::function<void()> task) { AsyncWorkers::get().submit(std::move(task)); }

void set_status(bsq_status* st, int code, const std::string& msg) {
  if (!st) return;
  st->code = code;
  snprintf(st->message, sizeof(st->message), "%s", msg.c_str());
}

std::string fmt_double(double v, unsigned precision, bool scientific) {
  char buf[400];
  const int n = fmt_double_buf(v, precision, scientific, buf, sizeof(buf));
  return std::string(buf, (size_t)(n > 0 ? n : 0));
}

// printf("%.*e") / printf("%.*f") digits through std::to_chars, which is specified to produce the
// same correctly rounded decimal form; NaN prints as NA.
int fmt_double_buf(double v, unsigned precision, bool scientific, char* buf, size_t len) {
  if (std::isnan(v)) {
    if (len < 3) return -1;
    buf[0] = 'N';
    buf[1] = 'A';
    buf[2] = 0;
    return 2;
  }
  const std::to_chars_result r =
      std::to_chars(buf, buf + len - 1, v, scientific ? std::chars_format::scientific : std::chars_format::fixed,
                    (int)precision);
  if (r.ec != std::errc()) return -1;
  int n = (int)(r.ptr - buf);
  // a 3-digit exponent with a leading zero is shortened (only non-glibc runtimes print one)
  if (scientific && n >= 3 && buf[n - 3] == '0') {
    buf[n - 3] = buf[n - 2];
    buf[n - 2] = buf[n - 1];
    n--;
  }
  buf[n] = 0;
  return n;
}

// printf("%.*g") through std::to_chars (general format with a precision is specified as %g)
std::string fmt_g(double v, int precision) {
  char buf[64];
  const std::to_chars_result r = std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::general, precision);
  if (r.ec != std::errc()) {
    snprintf(buf, sizeof(buf), "%.*g", precision, v);
    return buf;
  }
  return std::string(buf, (size_t)(r.ptr - buf));
}

double g6_round_trip(double v) {
  const double a = std::fabs(v);
  if (!std::isfinite(v) || !(a >= 1e-9 && a < 1e15)) return strtod(fmt_g(v).c_str(), nullptr);
  typedef unsigned __int128 u128;
  static const double kPow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                                    1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
  static const uint64_t kPow5[17] = {1ull,           5ull,           25ull,          125ull,         625ull,
                                     3125ull,        15625ull,       78125ull,       390625ull,      1953125ull,
                                     9765625ull,     48828125ull,    244140625ull,   1220703125ull,  6103515625ull,
                                     30517578125ull, 152587890625ull};
  // a = m * 2^be exactly
  int e2 = 0;
  const double fr = std::frexp(a, &e2);
  const uint64_t m = (uint64_t)std::ldexp(fr, 53);
  const int be = e2 - 53;
  // floor(log10(a)) is floor((e2 - 1) log10 2) or one more; a wrong guess is corrected below
  int d = (int)std::floor((double)(e2 - 1) * 0.30102999566398120);
  const int d1 = d + 1;
  if (d1 >= 0 ? a >= kPow10[d1] : a * kPow10[-d1] >= 1.0) d = d1;
  for (int tries = 0; tries < 4; tries++) {
    const int k = 5 - d;   // w = a * 10^k rounded, six digits when d is right
    // a * 10^k = N / D with N = m * 5^max(k,0) and D = 5^max(-k,0) * 2^sh, sh = -(be + k) > 0 here
    const int sh = -(be + k);
    if (k < -16 || k > 16 || sh <= 0 || sh >= 120) break;
    u128 w;
    if (k >= 0) {   // D is a power of two: shift and mask
      const u128 N = (u128)m * kPow5[k];
      w = N >> sh;
      const u128 r = N & ((((u128)1) << sh) - 1), h = ((u128)1) << (sh - 1);
      if (r > h || (r == h && (w & 1u))) w++;
    } else {
      const u128 D = ((u128)kPow5[-k]) << sh;
      w = (u128)m / D;
      const u128 r = (u128)m - w * D;
      if (2 * r > D || (2 * r == D && (w & 1u))) w++;
    }
    if (w >= 1000000u) {
      d++;
      continue;
    }
    if (w < 100000u) {
      d--;
      continue;
    }
    const double x = (double)(uint64_t)w;
    const double y = (k >= 0) ? x / kPow10[k] : x * kPow10[-k];
    return v < 0 ? -y : y;
  }
  return strtod(fmt_g(v).c_str(), nullptr);
}

static bool repeat_at(const char* s, int64_t len, int64_t pos, int64_t size, int64_t num) {
  const int64_t off = size * num;
  for (int64_t i = 0; i < size; i++) {
    if (pos + i >= len || pos + i + off >= len) return false;
    if (s[pos + i] != s[pos + i + off]) return false;
  }
  return true;
}

// SequenceTrims (calculate_trims.cpp:67-138) as a gather: each position's perfect-repeat extent
// m[p] (at least 1, capped at 255), then for every x the largest right trim x - p + 1 and left trim
// m[p] - (x - p) over the p in (x - 255, x] whose extent reaches x.  Both passes are split over host
// threads; every output byte is written by one thread.
void host_sequence_trims(const char* s, int64_t len, uint8_t* t) {
  std::vector<uint8_t> m((size_t)len);
  parallel_ranges(len, host_threads(), [&](int, int64_t p0, int64_t p1) {
    for (int64_t p = p0; p < p1; p++) {
      uint32_t mx = 1;
      for (int64_t size = 1; size <= 18; size++) {
        int64_t num = 1;
        while (num * size <= 255 && repeat_at(s, len, p, size, num)) num++;
        if (num > 1) mx = std::max<uint32_t>(mx, (uint32_t)(num * size));
      }
      m[(size_t)p] = (uint8_t)std::min<uint32_t>(mx, 255);
    }
  });
  parallel_ranges(len, host_threads(), [&](int, int64_t x0, int64_t x1) {
    for (int64_t x = x0; x < x1; x++) {
      uint32_t right = 0, left = 0;
      for (int64_t p = std::max<int64_t>(0, x - 254); p <= x; p++) {
        const uint32_t mp = m[(size_t)p], o = (uint32_t)(x - p);
        if (o < mp) {
          right = std::max(right, o + 1);
          left = std::max(left, mp - o);
        }
      }
      t[x] = (uint8_t)left;
      t[len + x] = (uint8_t)right;
    }
  });
}

static char comp(char c) {
  switch (c) {
    case 'A': return 'T';
    case 'C': return 'G';
    case 'G': return 'C';
    case 'T': return 'A';
  }
  return c;
}

void host_edge_trims(const std::string& seq, uint32_t& left, uint32_t& right) {
  left = right = 1;
  const int64_t n = (int64_t)seq.size();
  std::string rc(seq.rbegin(), seq.rend());
  for (char& c : rc) c = comp(c);
  for (int64_t size = 1; size <= 18; size++) {
    int64_t num = 1;
    while (repeat_at(seq.data(), n, 0, size, num)) num++;
    if (num > 1) left = std::max<uint32_t>(left, (uint32_t)(num * size));
    num = 1;
    while (repeat_at(rc.data(), n, 0, size, num)) num++;
    if (num > 1) right = std::max<uint32_t>(right, (uint32_t)(num * size));
  }
}

}  // namespace bsq
