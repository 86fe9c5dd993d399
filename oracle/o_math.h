/* Transcendentals of the oracle (test infrastructure).  The oracle calls the
   portable exp/log/pow of koral_amd/csrc/kd_math.h, the same functions the
   device calls, so a device run and its check round alike; they agree with
   the host libm that KORAL's build uses to within 0.5 ulp (correctly rounded
   except within about 2^-66 of a rounding boundary; tests/test_kd_math.py
   measures both).  The operation-counting build (oracle/flops) counts each
   call as one transcendental operation. */
#ifndef O_MATH_H
#define O_MATH_H
#ifdef O_FLOPCOUNT
#define O_POW pow
#define O_LOG log
#else
#include "kd_math.h"
#define O_POW kd_pow
#define O_LOG kd_log
#endif
#endif
