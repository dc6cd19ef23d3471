// bfa_internal.hpp -- host-side helpers shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include "../../include/blasfeo_amd.h"

namespace bfa {

// Records a message for blasfeo_amd_last_error() (thread-local) and returns code.
int set_error(int code, const char *fmt, ...);

// Checks the launch that was just enqueued.
int check_launch(const char *what);

// Number of compute units of the current device (cached per device).
int num_cus();
int lds_check(const void *fn, size_t dyn, const char *what);

// Device scratch owned by the library for one (device, stream) pair, at least
// `bytes` long.  Allocated on first use and replaced by a larger block when a
// larger size is asked for (the only allocation a batched entry may make);
// replaced blocks stay allocated until the process ends, so a HIP graph that
// captured an earlier block still finds it.  Launches on one stream are
// ordered, so reusing a block across calls on that stream is safe.
void *workspace(size_t bytes, hipStream_t stream);
// A second such block for the size plans of the leaf launches (the LDS-resident
// factorisations), which composite entries call while their own records sit in
// workspace().
void *plan_workspace(size_t bytes, hipStream_t stream);

// D = B * A^{-T} (trsm_rltn, alpha from the records) with the reciprocal
// diagonal of A read from sA->dA whatever its use_dA, and no cache update:
// the rows below the diagonal block of a trapezoid Cholesky (hint->max_n <= 256).
int trsm_rltn_given_dA(int ps, int batch, const blasfeo_amd_dtrsm_arg *args, const blasfeo_amd_hint *hint, hipStream_t st, const char *name);

// Launch-shape knobs (blasfeo_amd_tune_set); 0 means the built-in choice.
enum TuneKnob { TUNE_RIC_PER_CU = 0, TUNE_RIC_THREADS = 1, TUNE_GEMM_GRID_Y = 2, TUNE_POTRF_BW = 3, TUNE_TRSM_THIN = 4, TUNE_GEMM_WAVE = 5, TUNE_LDS_THREADS = 6 };
int tune_value(TuneKnob knob, int index);

} // namespace bfa
