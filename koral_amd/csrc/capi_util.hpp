// Error reporting shared by the C-ABI entry points.
#pragma once
#include <hip/hip_runtime.h>
#include <stdio.h>
#include "koral_hip.h"

namespace kd {

char *last_error_buf();

inline int set_error(int code, const char *msg)
{
  snprintf(last_error_buf(), 512, "%s", msg);
  return code;
}

inline int check_params(const koral_params *P)
{
  if (!P) return set_error(KORAL_E_ARG, "null koral_params");
  int nv = 6 + 3 * (P->mag ? 1 : 0) + 4 * (P->rad ? 1 : 0);
  if (P->nv != nv) return set_error(KORAL_E_ARG, "koral_params.nv inconsistent with mag/rad");
  return 0;
}

inline int check_hip(hipError_t e, const char *where)
{
  if (e == hipSuccess) return 0;
  snprintf(last_error_buf(), 512, "%s: %s", where, hipGetErrorString(e));
  return KORAL_E_HIP;
}

inline int check_launch(const char *where) { return check_hip(hipGetLastError(), where); }

}  // namespace kd
