// prof.h -- optional per-launch timing of the hot kernels with HIP events
// recorded on the launch stream (off by default; enabled by
// scuffx_profile).  Each timed launch also carries its algorithmic work
// (flops or bytes) so that callers can form achieved rates.
#pragma once

#include <hip/hip_runtime.h>

namespace scuffx {

enum ProfKernel {
  PK_GEMM_C128 = 0, PK_GEMM_C64, PK_PANEL_C128, PK_PANEL_C64, PK_TRSM_C128, PK_TRSM_C64,
  PK_FAR, PK_NEAR, PK_SELF, PK_STAMP, PK_FIPPI, PK_RESID, PK_GRAD, PK_GEMM_F64, PK_PANEL_F64, PK_COUNT
};

bool prof_enabled();
void prof_enable(bool on);
// returns a token to pass to prof_end, or -1 when profiling is off
int prof_begin(hipStream_t s, int kernel);
void prof_end(hipStream_t s, int token, double work);
// synchronizes the recorded events and folds them into the totals
void prof_collect();
// totals since the last reset: milliseconds, work, launches
void prof_read(int kernel, double *ms, double *work, long *launches);
void prof_reset();
const char *prof_name(int kernel);

}  // namespace scuffx
