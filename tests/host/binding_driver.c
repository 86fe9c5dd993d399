/*
 * Test driver of the KORAL-signature binding (koral_amd/binding/koral_ops.c),
 * built with gcc against include/koral_hip.h and linked with libkoral_hip.so:
 * it plays ko.c's part.  It reads a grid set-up (the host check's input
 * format), initialises the state as KORAL does (u = p2u(p), ghosts, dt from
 * the wave speeds), then runs NSTEPS RK2IMEX steps with ko.c's stage algebra
 * (ko.c:444-570) on the host arrays, calling op_implicit, calc_u2p,
 * do_correct and op_explicit through KORAL's own signatures.  Output: p, u,
 * flags, then per step tstepdenmax and six zero counters.
 *
 * usage: binding_driver IN OUT NSTEPS
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "koral_ops.h"

ldouble *p, *u;
int *cellflag;
ldouble tstepdenmax;

static void rd(FILE *f, void *x, size_t n, const char *what)
{
  if (fread(x, 1, n, f) != n) {
    fprintf(stderr, "binding_driver: short read (%s)\n", what);
    exit(2);
  }
}

static double *rd_doubles(FILE *f, long n, const char *what)
{
  double *x = malloc(sizeof(double) * (n > 0 ? n : 1));
  rd(f, x, sizeof(double) * n, what);
  return x;
}

int main(int argc, char **argv)
{
  if (argc < 4) {
    fprintf(stderr, "usage: binding_driver IN OUT NSTEPS\n");
    return 2;
  }
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 2;
  int32_t hdr[12];
  rd(f, hdr, sizeof hdr, "header");
  if (hdr[0] != 0x4b4f5241 || hdr[11] != (int)sizeof(koral_params)) {
    fprintf(stderr, "binding_driver: bad header\n");
    return 2;
  }
  koral_params P;
  rd(f, &P, sizeof P, "params");
  koral_grid_desc D;
  memset(&D, 0, sizeof D);
  D.nx = hdr[1];
  D.ny = hdr[2];
  D.nz = hdr[3];
  D.ng = hdr[4];
  for (int k = 0; k < 6; k++) D.bc[k] = hdr[5 + k];
  const int NX = D.nx, NY = D.ny, NZ = D.nz, NG = D.ng, NV = P.nv;
  const int ngx = NG, ngy = NY > 1 ? NG : 0, ngz = NZ > 1 ? NG : 0;
  const int SX = NX + 2 * ngx, SY = NY + 2 * ngy, SZ = NZ + 2 * ngz;
  const int nd[3] = {NX, NY, NZ};
  for (int d = 0; d < 3; d++) D.xb[d] = rd_doubles(f, nd[d] + 2 * NG + 1, "xb");
  const long npl = (long)SY * SX, nplx = (long)SY * (SX + 1), nply = (long)(SY + 1) * SX;
  D.gc = rd_doubles(f, 20 * npl, "gc");
  D.Gc = rd_doubles(f, 20 * npl, "Gc");
  D.kr = rd_doubles(f, 64 * npl, "kr");
  D.gfx = rd_doubles(f, 20 * nplx, "gfx");
  D.Gfx = rd_doubles(f, 20 * nplx, "Gfx");
  D.gfy = rd_doubles(f, 20 * nply, "gfy");
  D.Gfy = rd_doubles(f, 20 * nply, "Gfy");
  const long NC = (long)SX * SY * SZ;
  p = rd_doubles(f, NV * NC, "p");
  fclose(f);
  u = calloc(NV * NC, sizeof(double));
  cellflag = calloc(7 * NC, sizeof(int));
  if (koral_ops_bind(&P, &D)) {
    fprintf(stderr, "binding_driver: bind: %s\n", koral_hip_last_error());
    return 1;
  }
  /* initial state as ko.c leaves it: u = p2u(p) everywhere, ghosts filled,
     the first dt from the wave speeds */
  koral_hip_grid *KG = koral_ops_grid();
  double tsd0 = 0.;
  if (koral_hip_grid_upload(KG, p, NULL, cellflag) || koral_hip_grid_p2u(KG) || koral_hip_set_bc(KG) ||
      koral_hip_tstepden(KG, &tsd0) || koral_hip_grid_download(KG, p, u, cellflag)) {
    fprintf(stderr, "binding_driver: init: %s\n", koral_hip_last_error());
    return 1;
  }
  const int nsteps = atoi(argv[3]);
  double *ut0 = malloc(sizeof(double) * NV * NC), *ut1 = malloc(sizeof(double) * NV * NC);
  double *ut2 = malloc(sizeof(double) * NV * NC), *uforget = malloc(sizeof(double) * NV * NC);
  double *dut1 = malloc(sizeof(double) * NV * NC), *dut2 = malloc(sizeof(double) * NV * NC);
  double *drt1 = malloc(sizeof(double) * NV * NC), *drt2 = malloc(sizeof(double) * NV * NC);
  double *tsds = malloc(sizeof(double) * (nsteps > 0 ? nsteps : 1));
  ldouble t = 0., dt = 1. / tsd0;
  const ldouble gamma = 1. - 1. / sqrt(2.);
#define DOM for (int iz = 0; iz < NZ; iz++) for (int iy = 0; iy < NY; iy++) for (int ix = 0; ix < NX; ix++)
#define C(a) a[(((long)(iz + ngz) * SY + iy + ngy) * SX + ix + ngx) * NV + iv]
  for (int k = 0; k < nsteps; k++) {
    tstepdenmax = 0.;
    /* ko.c:444-570 */
    DOM for (int iv = 0; iv < NV; iv++) C(ut0) = 1. * C(u);
    op_implicit(t, dt * gamma);
    DOM for (int iv = 0; iv < NV; iv++) C(drt1) = (1. / (dt * gamma)) * C(u) + (-1. / (dt * gamma)) * C(ut0);
    DOM for (int iv = 0; iv < NV; iv++) C(ut1) = 1. * C(u);
    calc_u2p();
    do_correct();
    op_explicit(t, dt);
    DOM for (int iv = 0; iv < NV; iv++) C(dut1) = (1. / (dt)) * C(u) + (-1. / (dt)) * C(ut1);
    DOM for (int iv = 0; iv < NV; iv++) C(u) = C(ut0) + (dt) * C(dut1) + (dt * (1. - 2. * gamma)) * C(drt1);
    DOM for (int iv = 0; iv < NV; iv++) C(uforget) = 1. * C(u);
    calc_u2p();
    do_correct();
    op_implicit(t, gamma * dt);
    DOM for (int iv = 0; iv < NV; iv++) C(drt2) = (1. / (dt * gamma)) * C(u) + (-1. / (dt * gamma)) * C(uforget);
    DOM for (int iv = 0; iv < NV; iv++) C(ut2) = 1. * C(u);
    calc_u2p();
    do_correct();
    op_explicit(t, dt);
    DOM for (int iv = 0; iv < NV; iv++) C(dut2) = (1. / (dt)) * C(u) + (-1. / (dt)) * C(ut2);
    DOM for (int iv = 0; iv < NV; iv++) C(u) = C(ut0) + (dt / 2.) * C(dut1) + (dt / 2.) * C(dut2);
    DOM for (int iv = 0; iv < NV; iv++) C(u) = C(u) + (dt / 2.) * C(drt1) + (dt / 2.) * C(drt2);
    calc_u2p();
    do_correct();
    tsds[k] = tstepdenmax;
    t += dt;
    dt = 1. / tstepdenmax; /* ko.c:396 */
    fprintf(stderr, "binding_driver: step %d done, tstepdenmax %.17e\n", k, tstepdenmax);
  }
#undef DOM
#undef C
  koral_ops_unbind();
  FILE *o = fopen(argv[2], "wb");
  if (!o) return 2;
  fwrite(p, sizeof(double), NV * NC, o);
  fwrite(u, sizeof(double), NV * NC, o);
  fwrite(cellflag, sizeof(int), 7 * NC, o);
  const int zeros[6] = {0, 0, 0, 0, 0, 0};
  for (int k = 0; k < nsteps; k++) {
    fwrite(&tsds[k], sizeof(double), 1, o);
    fwrite(zeros, sizeof(int), 6, o);
  }
  fclose(o);
  return 0;
}
