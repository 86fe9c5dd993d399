// Function qualifiers of the per-point, per-face and per-cell device code.
//
// The kernels are HIP for gfx950 only.  Defining KD_HOST_CHECK compiles the
// same functions as plain inline C++ so that a host build can drive them
// under address sanitising and bounds checks (tests/host/grid_check.cpp);
// nothing in the library is ever built or run that way.
#pragma once
#ifdef KD_HOST_CHECK
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#define KD_INL inline
#define KD_FUN inline
#define KD_HD inline
#else
#include <hip/hip_runtime.h>
#define KD_INL __device__ __forceinline__
#define KD_FUN __device__
#define KD_HD __host__ __device__ inline
#endif

// Index checks of the grid addressing: a coordinate outside the box an index
// function serves stops a host-check build with a message naming the
// function and the coordinates.  The device build compiles them to nothing.
#ifdef KD_HOST_CHECK
inline void kd_index_fail(const char *what, long a, long b, long c)
{
  fprintf(stderr, "INDEX CHECK FAILED: %s (%ld, %ld, %ld)\n", what, a, b, c);
  fflush(stderr);
  abort();
}
#define KD_CHECK(cond, what, a, b, c)                                    \
  do {                                                                    \
    if (!(cond)) ::kd_index_fail(what, (long)(a), (long)(b), (long)(c)); \
  } while (0)
#else
#define KD_CHECK(cond, what, a, b, c) ((void)0)
#endif

// Compiler-only ordering point: no memory access is moved across it, which
// bounds how many loads the scheduler hoists (and keeps live) at once.  It
// emits no instruction.
#ifdef KD_HOST_CHECK
inline void compiler_fence() { asm volatile("" ::: "memory"); }
#else
__device__ __forceinline__ void compiler_fence() { asm volatile("" ::: "memory"); }
#endif
