// ws.cpp -- per-thread, per-device grow-only workspaces.
#include "ws.h"

#include <hip/hip_runtime.h>

#include <map>

namespace scuffx {

namespace {
struct Buf { void *p = nullptr; size_t cap = 0; };
struct Pool {
  std::map<int, Buf[WS_COUNT]> dev;
  ~Pool()
  {
    for (auto &kv : dev)
      for (auto &b : kv.second) if (b.p) hipFree(b.p);
  }
};
thread_local Pool g_pool;
}  // namespace

void *ws_get(int slot, size_t bytes)
{
  if (slot < 0 || slot >= WS_COUNT) return nullptr;
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess) return nullptr;
  Buf &b = g_pool.dev[d][slot];
  if (b.p && b.cap >= bytes) return b.p;
  if (b.p) { hipDeviceSynchronize(); hipFree(b.p); b.p = nullptr; b.cap = 0; }
  size_t want = bytes < 256 ? 256 : bytes;
  want += want / 8;   // headroom for slightly larger later calls
  if (hipMalloc(&b.p, want) != hipSuccess) { b.p = nullptr; return nullptr; }
  b.cap = want;
  return b.p;
}

}  // namespace scuffx
