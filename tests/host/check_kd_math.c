/* Host driver of kd_math.h (koral_amd/csrc) built as C: reads records of
   (op, x, y) from stdin, op 0 exp(x), 1 log(x), 2 pow(x, y), and writes
   kd_* and the host libm result per record to stdout, all as raw doubles. */
#include <stdio.h>
#include <math.h>
#include "kd_math.h"

int main(void)
{
  double rec[3];
  while (fread(rec, sizeof(double), 3, stdin) == 3) {
    int op = (int)rec[0];
    double x = rec[1], y = rec[2], a = 0., b = 0.;
    if (op == 0) {
      a = kd_exp(x);
      b = exp(x);
    } else if (op == 1) {
      a = kd_log(x);
      b = log(x);
    } else {
      a = kd_pow(x, y);
      b = pow(x, y);
    }
    double out[2] = {a, b};
    fwrite(out, sizeof(double), 2, stdout);
  }
  return 0;
}
