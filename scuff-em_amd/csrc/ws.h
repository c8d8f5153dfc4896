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
  WS_GMRES,        // GMRES refinement: Krylov basis and vectors
  WS_SMALL,        // reductions with a host result (log det, trace, norms)
  WS_STAGE,        // staging of blocks (real blocks, LU inverse, unpacked storage)
  WS_COUNT
};

// device pointer of at least `bytes` bytes for `slot` on the current device
// (nullptr if the allocation fails or the library was shut down)
void *ws_get(int slot, size_t bytes);

// frees every workspace of every thread and device (at shutdown)
void ws_release_all();

// after scuffx_shutdown the library makes no further HIP calls: device
// memory still held by live objects is left to the process exit
bool hip_down();
void set_hip_down();

}  // namespace scuffx
