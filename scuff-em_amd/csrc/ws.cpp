// ws.cpp -- grow-only device workspaces per host thread and device, and
// the library's shutdown flag.
#include "ws.h"

#include <hip/hip_runtime.h>

#include <atomic>
#include <map>
#include <mutex>
#include <thread>
#include <utility>

namespace scuffx {

namespace {
struct Buf { void *p = nullptr; size_t cap = 0; };
// every pool of every thread, so that shutdown can release them all while
// the HIP runtime is still up (thread-exit destructors may run after it)
std::mutex g_mu;
std::map<std::pair<std::thread::id, int>, Buf[WS_COUNT]> *g_pools = nullptr;
std::atomic<bool> g_down{false};
}  // namespace

bool hip_down() { return g_down.load(std::memory_order_acquire); }

void *ws_get(int slot, size_t bytes)
{
  if (slot < 0 || slot >= WS_COUNT || hip_down()) return nullptr;
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_pools) g_pools = new std::map<std::pair<std::thread::id, int>, Buf[WS_COUNT]>();
  Buf &b = (*g_pools)[{std::this_thread::get_id(), d}][slot];
  if (b.p && b.cap >= bytes) return b.p;
  if (b.p) { hipDeviceSynchronize(); hipFree(b.p); b.p = nullptr; b.cap = 0; }
  size_t want = bytes < 256 ? 256 : bytes;
  want += want / 8;   // headroom for slightly larger later calls
  if (hipMalloc(&b.p, want) != hipSuccess) { b.p = nullptr; return nullptr; }
  b.cap = want;
  return b.p;
}

void ws_release_all()
{
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_pools) return;
  for (auto &kv : *g_pools) {
    hipSetDevice(kv.first.second);
    hipDeviceSynchronize();
    for (auto &b : kv.second)
      if (b.p) { hipFree(b.p); b.p = nullptr; b.cap = 0; }
  }
  g_pools->clear();
}

void set_hip_down() { g_down.store(true, std::memory_order_release); }

}  // namespace scuffx
