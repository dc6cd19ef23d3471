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
// `bytes` long.  Allocated on first use and grown when a larger size is asked
// for (the only allocation a batched entry may make); launches on one stream
// are ordered, so reusing it across calls on that stream is safe.
void *workspace(size_t bytes, hipStream_t stream);

// Launch-shape knobs (blasfeo_amd_tune_set); 0 means the built-in choice.
enum TuneKnob { TUNE_RIC_PER_CU = 0, TUNE_RIC_THREADS = 1, TUNE_GEMM_GRID_Y = 2, TUNE_POTRF_BW = 3, TUNE_TRSM_THIN = 4 };
int tune_value(TuneKnob knob, int index);

} // namespace bfa
