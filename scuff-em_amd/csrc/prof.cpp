// prof.cpp -- event-pair timing of individual kernel launches.
#include "prof.h"

#include <cstring>
#include <mutex>
#include <vector>

namespace scuffx {

namespace {
struct Rec { hipEvent_t a, b; int kernel; double work; };
std::mutex g_mu;
bool g_on = false;
std::vector<Rec> g_open;
std::vector<hipEvent_t> g_pool;
double g_ms[PK_COUNT], g_work[PK_COUNT];
long g_n[PK_COUNT];
const char *g_names[PK_COUNT] = {"gemm_c128", "gemm_c64", "panel_c128", "panel_c64", "trsm_c128", "trsm_c64",
                                 "far", "near", "self", "stamp", "fippi", "resid", "grad", "gemm_f64", "panel_f64"};

hipEvent_t take()
{
  if (!g_pool.empty()) { hipEvent_t e = g_pool.back(); g_pool.pop_back(); return e; }
  // timing only: no system-scope fence (cache write-back and invalidation)
  // when the event is recorded, which would slow the work it brackets
  hipEvent_t e = nullptr;
  hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
  return e;
}
}  // namespace

bool prof_enabled() { return g_on; }

void prof_enable(bool on)
{
  std::lock_guard<std::mutex> l(g_mu);
  g_on = on;
}

int prof_begin(hipStream_t s, int kernel)
{
  if (!g_on) return -1;
  // the block-row kernels of the LU (one to two small launches per panel)
  // are not timed: an event pair costs ~3 us, about what such a launch
  // takes, and they are never the largest kernel (their time is in the
  // rocprofv3 summaries)
  if (kernel == PK_TRSM_C128 || kernel == PK_TRSM_C64) return -1;
  std::lock_guard<std::mutex> l(g_mu);
  Rec r;
  r.a = take(); r.b = take(); r.kernel = kernel; r.work = 0;
  hipEventRecord(r.a, s);
  g_open.push_back(r);
  return (int)g_open.size() - 1;
}

void prof_end(hipStream_t s, int token, double work)
{
  if (token < 0) return;
  std::lock_guard<std::mutex> l(g_mu);
  if (token >= (int)g_open.size()) return;
  g_open[token].work = work;
  hipEventRecord(g_open[token].b, s);
}

void prof_collect()
{
  std::lock_guard<std::mutex> l(g_mu);
  for (auto &r : g_open) {
    hipEventSynchronize(r.b);
    float ms = 0;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      g_ms[r.kernel] += ms;
      g_work[r.kernel] += r.work;
      g_n[r.kernel] += 1;
    }
    g_pool.push_back(r.a);
    g_pool.push_back(r.b);
  }
  g_open.clear();
}

void prof_read(int kernel, double *ms, double *work, long *launches)
{
  std::lock_guard<std::mutex> l(g_mu);
  if (kernel < 0 || kernel >= PK_COUNT) return;
  if (ms) *ms = g_ms[kernel];
  if (work) *work = g_work[kernel];
  if (launches) *launches = g_n[kernel];
}

void prof_reset()
{
  std::lock_guard<std::mutex> l(g_mu);
  std::memset(g_ms, 0, sizeof g_ms);
  std::memset(g_work, 0, sizeof g_work);
  std::memset(g_n, 0, sizeof g_n);
}

const char *prof_name(int kernel) { return (kernel >= 0 && kernel < PK_COUNT) ? g_names[kernel] : ""; }

}  // namespace scuffx
