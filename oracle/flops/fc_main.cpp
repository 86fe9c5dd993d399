// Average operation counts per implicit solve and per u2p inversion of the
// oracle on a batch read from a file written by koral_amd (seeded states),
// and per zone-update of each operator of one RK2IMEX grid step on a grid
// set-up read from a file in the host check's input format.
#include <stdio.h>
#include "cd.hpp"
#define double cd
#include "../oracle.h"
#undef double
FlopCount g_flops;

static int rd(FILE *f, void *p, size_t n) { return fread(p, 1, n, f) == n ? 0 : -1; }
int fc_grid_counts(FILE *f);

int main(int argc, char **argv)
{
  if (argc < 3) { fprintf(stderr, "usage: flopcount implicit|u2p batch.bin | grid input.bin\n"); return 2; }
  FILE *f = fopen(argv[2], "rb");
  if (!f) return 2;
  if (argv[1][0] == 'g') return fc_grid_counts(f);
  oracle_params P;
  long n = 0;
  if (rd(f, &P, sizeof P) || rd(f, &n, sizeof n)) return 2;
  int nv = P.nv;
  cd *pp = (cd *)malloc(sizeof(cd) * nv * n), *uu = (cd *)malloc(sizeof(cd) * nv * n);
  cd *g20 = (cd *)malloc(sizeof(cd) * 20 * n), *G20 = (cd *)malloc(sizeof(cd) * 20 * n), *dt = (cd *)malloc(sizeof(cd) * n);
  if (rd(f, pp, sizeof(cd) * nv * n) || rd(f, uu, sizeof(cd) * nv * n) || rd(f, g20, sizeof(cd) * 20 * n) ||
      rd(f, G20, sizeof(cd) * 20 * n) || rd(f, dt, sizeof(cd) * n)) return 2;
  fclose(f);
  long ok = 0, iters = 0;
  memset(&g_flops, 0, sizeof g_flops);
  for (long i = 0; i < n; i++) {
    o_geom g;
    o_fill_geom(&g, g20 + 20 * i, G20 + 20 * i);
    FlopCount before = g_flops;
    if (argv[1][0] == 'i') {
      int flags[7] = {0};
      o_imp_info inf;
      int r = o_solve_implicit_lab(&P, uu + nv * i, pp + nv * i, &g, dt[i], 0., flags, &inf);
      ok += (r == 0);
      iters += inf.total_iters;
    } else {
      int corr[3] = {0, 0, 0}, fix[2] = {0, 0};
      o_u2p(&P, uu + nv * i, pp + nv * i, &g, corr, fix);
    }
    (void)before;
  }
  double N = (double)n;
  printf("{\"n\": %ld, \"add\": %.1f, \"mul\": %.1f, \"div\": %.1f, \"sqrt\": %.1f, \"trans\": %.1f, "
         "\"flops\": %.1f, \"solved\": %ld, \"iters\": %.6f}\n",
         n, g_flops.add / N, g_flops.mul / N, g_flops.div / N, g_flops.sqrt / N, g_flops.trans / N,
         (g_flops.add + g_flops.mul + g_flops.div + g_flops.sqrt + g_flops.trans) / N, ok, iters / N);
  return 0;
}
