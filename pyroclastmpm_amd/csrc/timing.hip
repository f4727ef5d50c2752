// timing.hip -- per-thread kernel timing with HIP events (pmpm_timing_*).
//
// When enabled on a thread, every kernel the library launches from that
// thread (PMPM_LAUNCH) is bracketed by a pair of HIP events recorded on the
// launch stream. Reading waits for the events, sums the elapsed times per
// kernel name and returns the events to a pool. Off by default: a launch then
// costs one thread-local flag test.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "context.hpp"

namespace pmpm {
namespace {

struct Timed {
  std::string name;
  hipEvent_t a, b;
};
struct TimingState {
  bool on = false;
  std::vector<Timed> pending;
  std::vector<hipEvent_t> pool;
  hipEvent_t take() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
  ~TimingState() {
    for (auto &t : pending) { (void)hipEventDestroy(t.a); (void)hipEventDestroy(t.b); }
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};
thread_local TimingState t_timing;

}  // namespace

bool timing_enabled() { return t_timing.on; }

std::string kname(const char *base, std::initializer_list<long> targs) {
  std::string s(base);
  if (targs.size() == 0) return s;
  s += '<';
  bool first = true;
  for (long v : targs) {
    if (!first) s += ", ";
    s += std::to_string(v);
    first = false;
  }
  s += '>';
  return s;
}

void TimedRegion::start(hipStream_t s, std::string name) {
  hipEvent_t a = t_timing.take(), b = t_timing.take();
  if (!a || !b) return;
  if (hipEventRecord(a, s) != hipSuccess) {
    t_timing.pool.push_back(a);
    t_timing.pool.push_back(b);
    return;
  }
  on_ = true;
  s_ = s;
  a_ = a;
  b_ = b;
  name_ = std::move(name);
}

void TimedRegion::finish() {
  (void)hipEventRecord(b_, s_);
  t_timing.pending.push_back(Timed{std::move(name_), a_, b_});
}

}  // namespace pmpm

using namespace pmpm;

extern "C" int pmpm_timing_enable(int32_t enable) {
  t_timing.on = enable != 0;
  return PMPM_OK;
}

extern "C" int pmpm_timing_read(char *names, int32_t names_len, int32_t *counts, float *ms, int32_t max_kernels,
                                int32_t *num_kernels) {
  if (max_kernels < 0 || (max_kernels > 0 && (!counts || !ms))) return PMPM_ERR_INVALID;
  std::vector<std::string> order;
  std::vector<int32_t> cnt;
  std::vector<double> tot;
  int rc = PMPM_OK;
  for (auto &t : t_timing.pending) {
    float e = 0.f;
    if (rc == PMPM_OK) {
      hipError_t st = hipEventSynchronize(t.b);
      if (st == hipSuccess) st = hipEventElapsedTime(&e, t.a, t.b);
      if (st != hipSuccess) {
        note_hip_error(st, "hipEventElapsedTime");
        rc = PMPM_ERR_HIP;
      }
    }
    size_t k = 0;
    while (k < order.size() && order[k] != t.name) ++k;
    if (k == order.size()) {
      order.push_back(t.name);
      cnt.push_back(0);
      tot.push_back(0.0);
    }
    cnt[k] += 1;
    tot[k] += e;
    t_timing.pool.push_back(t.a);
    t_timing.pool.push_back(t.b);
  }
  t_timing.pending.clear();
  if (rc != PMPM_OK) return rc;
  const int32_t nk = (int32_t)order.size();
  if (num_kernels) *num_kernels = nk;
  std::string joined;
  for (int32_t k = 0; k < nk; ++k) {
    if (k < max_kernels) {
      counts[k] = cnt[k];
      ms[k] = (float)tot[k];
    }
    if (k) joined += '\n';
    joined += order[k];
  }
  if (names && names_len > 0) {
    const size_t m = joined.size() < (size_t)(names_len - 1) ? joined.size() : (size_t)(names_len - 1);
    joined.copy(names, m);
    names[m] = '\0';
  }
  return PMPM_OK;
}
