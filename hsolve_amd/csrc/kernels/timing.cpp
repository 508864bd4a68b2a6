#include "timing.hpp"

#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace hsg {

namespace {

struct Record {
  hipEvent_t start = nullptr, stop = nullptr;
  int phase = 0;
  int64_t launches = 0;
  double work = 0.0;
  bool closed = false;
};

std::mutex g_mu;
int g_level = 0;
std::vector<Record> g_records;

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void clear_locked() {
  for (Record& r : g_records) {
    if (r.start) (void)hipEventDestroy(r.start);
    if (r.stop) (void)hipEventDestroy(r.stop);
  }
  g_records.clear();
}

}  // namespace

void timing_enable(int level) {
  std::lock_guard<std::mutex> lk(g_mu);
  clear_locked();
  g_level = level;
}

int timing_level() {
  std::lock_guard<std::mutex> lk(g_mu);
  return g_level;
}

int64_t timing_begin(int phase, hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_level == 0 || (phase == kPhaseUpdate && g_level < 2)) return -1;
  Record r;
  r.phase = phase;
  check(hipEventCreate(&r.start), "hipEventCreate (timing)");
  check(hipEventCreate(&r.stop), "hipEventCreate (timing)");
  check(hipEventRecord(r.start, st), "hipEventRecord (timing)");
  g_records.push_back(r);
  return (int64_t)g_records.size() - 1;
}

void timing_end(int64_t token, hipStream_t st, int64_t launches, double work) {
  if (token < 0) return;
  std::lock_guard<std::mutex> lk(g_mu);
  if (token >= (int64_t)g_records.size()) return;  // cleared in between
  Record& r = g_records[token];
  check(hipEventRecord(r.stop, st), "hipEventRecord (timing)");
  r.launches = launches;
  r.work = work;
  r.closed = true;
}

PhaseTotals timing_read(int phase) {
  std::lock_guard<std::mutex> lk(g_mu);
  PhaseTotals t;
  for (Record& r : g_records) {
    if (r.phase != phase || !r.closed) continue;
    check(hipEventSynchronize(r.stop), "hipEventSynchronize (timing)");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, r.start, r.stop), "hipEventElapsedTime (timing)");
    t.ms += ms;
    t.launches += r.launches;
    t.work += r.work;
    t.calls += 1;
  }
  return t;
}

}  // namespace hsg
