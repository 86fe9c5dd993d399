/* Oracle: grid-level operators (explicit step, u2p sweep, fixups, boundaries).
   Test infrastructure only. */
#include "oracle.h"
