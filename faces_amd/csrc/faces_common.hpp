// faces_common.hpp -- host-side helpers shared by the library's translation
// units: error reporting and the compile-time dispatch over element size n.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/faces_amd.h"

namespace faces {

// Records a failure for faces_last_error() and returns the status.
int fail(int status, const char *fmt, ...);

// Maps a HIP error to FACES_ERR_HIP (or FACES_ERR_NOMEM) with its text.
int hip_fail(hipError_t e, const char *what);

// Number of compute units of the current device (cached per device).
int device_cus();
int device_lds_per_cu();

// The current device's stream for the library's set-up work (non-blocking,
// created on first use).
hipError_t setup_stream(hipStream_t *out);

// The staged step (pack into receive slots, interior sums, surface sums) of
// one rank on a message arena of arena_bytes the caller provides; builds its
// plan on the host stack and allocates nothing (faces_plan.hip).  `who` names
// the caller in the error text.
int share_staged_on_arena(double *u, int rank, int lx, int ly, int lz, int mx, int my, int mz, int n,
                          int boundary, long plane_stride, int compute, double *arena, size_t arena_bytes,
                          const char *who, hipStream_t s);

}  // namespace faces

#define FACES_HIP(call)                                          \
  do {                                                           \
    hipError_t faces_e_ = (call);                                \
    if (faces_e_ != hipSuccess) return faces::hip_fail(faces_e_, #call); \
  } while (0)

// Element sizes with a compile-time specialisation; others run the generic
// (runtime n) instantiation N = 0.
#define FACES_DISPATCH_N(nval, MACRO) \
  switch (nval) {                     \
    case 2: MACRO(2); break;          \
    case 3: MACRO(3); break;          \
    case 4: MACRO(4); break;          \
    case 5: MACRO(5); break;          \
    case 6: MACRO(6); break;          \
    case 7: MACRO(7); break;          \
    case 8: MACRO(8); break;          \
    case 9: MACRO(9); break;          \
    case 10: MACRO(10); break;        \
    case 11: MACRO(11); break;        \
    case 12: MACRO(12); break;        \
    default: MACRO(0); break;         \
  }
