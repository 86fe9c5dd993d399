// Function qualifiers of the per-point, per-face and per-cell device code.
//
// The kernels are HIP for gfx950 only.  Defining KD_HOST_CHECK compiles the
// same functions as plain inline C++ so that a host build can drive them
// under address sanitising and bounds checks while a kernel is debugged;
// nothing in the library is ever built or run that way.
#pragma once
#ifdef KD_HOST_CHECK
#include <math.h>
#define KD_INL inline
#define KD_FUN inline
#define KD_HD inline
#else
#include <hip/hip_runtime.h>
#define KD_INL __device__ __forceinline__
#define KD_FUN __device__
#define KD_HD __host__ __device__ inline
#endif

// Compiler-only ordering point: no memory access is moved across it, which
// bounds how many loads the scheduler hoists (and keeps live) at once.  It
// emits no instruction.
#ifdef KD_HOST_CHECK
inline void compiler_fence() { asm volatile("" ::: "memory"); }
#else
__device__ __forceinline__ void compiler_fence() { asm volatile("" ::: "memory"); }
#endif
