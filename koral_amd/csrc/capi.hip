// Library-level C-ABI entry points (version, error text, struct size).
#include "koral_hip.h"
#include "capi_util.hpp"

namespace kd {
char *last_error_buf()
{
  static thread_local char buf[512] = "";
  return buf;
}
}  // namespace kd

extern "C" {
const char *koral_hip_last_error(void) { return kd::last_error_buf(); }
int koral_hip_version(void) { return 1; }
int koral_hip_params_size(void) { return (int)sizeof(koral_params); }
}
