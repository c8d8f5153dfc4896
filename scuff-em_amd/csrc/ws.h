// ws.h -- grow-only device workspaces, one set per host thread and device,
// so that the LU, the solves and the refinement do not allocate on every
// call.  A slot is reused by the next call that asks for it; two calls that
// overlap on one thread and device must not use the same slot (the library
// never does: each slot has a single owner).
#pragma once

#include <cstddef>

namespace scuffx {

enum WsSlot {
  WS_PANEL = 0,    // LU panel scratch (barrier words, candidates, plan)
  WS_SOLVE_W,      // triangular solve: intermediate vector(s)
  WS_SOLVE_INVL,   // triangular solve: inverses of the diagonal blocks of L
  WS_SOLVE_INVU,   // ... and of U
  WS_REFINE,       // refinement: residual, partial sums, complex64 residual, scalar
  WS_PERM,         // triangular solve: composed row permutation of ipiv
  WS_PERM_TMP,     // ... and the permuted right-hand sides (many-column solves)
  WS_CHAIN,        // one-launch triangular solves: error word and block flags
  WS_COUNT
};

// device pointer of at least `bytes` bytes for `slot` on the current device
// (nullptr if the allocation fails)
void *ws_get(int slot, size_t bytes);

}  // namespace scuffx
