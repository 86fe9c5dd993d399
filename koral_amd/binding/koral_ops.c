/*
 * KORAL's operators (op_explicit, op_implicit, calc_u2p, do_correct) on the
 * MI355X through libkoral_hip.so, in parity mode: KORAL's host arrays are the
 * state; each operator uploads p, u and cellflag, runs on the device and
 * downloads them again.  Errors stop the run as KORAL's my_err does.
 */
#include <stdio.h>
#include <stdlib.h>
#include "koral_ops.h"

static koral_hip_grid *KG;

static void fail(const char *what)
{
  fprintf(stderr, "koral_ops: %s: %s\n", what, koral_hip_last_error());
  exit(-1);
}

int koral_ops_bind(const koral_params *P, const koral_grid_desc *D)
{
  if (koral_hip_params_size() != (int)sizeof(koral_params)) {
    fprintf(stderr, "koral_ops: koral_params size differs from the library's\n");
    return -1;
  }
  return koral_hip_grid_create(P, D, &KG);
}

void koral_ops_unbind(void)
{
  koral_hip_grid_destroy(KG);
  KG = NULL;
}

koral_hip_grid *koral_ops_grid(void) { return KG; }

static void state_in(void)
{
  if (koral_hip_grid_upload(KG, p, u, cellflag)) fail("upload");
}

static void state_out(void)
{
  if (koral_hip_grid_download(KG, p, u, cellflag)) fail("download");
}

/* finite.c:647-1234: ghosts, wave speeds, fluxes, flux-CT, update;
   tstepdenmax keeps the largest time-step density seen (finite.c:296-398) */
int op_explicit(ldouble t, ldouble dt)
{
  double tsd = 0.;
  state_in();
  if (koral_hip_op_explicit(KG, t, dt, &tsd)) fail("op_explicit");
  if (tsd > tstepdenmax) tstepdenmax = tsd;
  state_out();
  return 0;
}

/* finite.c:1242-1292: the implicit radiation source of every cell */
int op_implicit(ldouble t, ldouble dt)
{
  state_in();
  if (koral_hip_op_implicit(KG, t, dt)) fail("op_implicit");
  state_out();
  return 0;
}

/* finite.c:502-561: u2p with floors and fixups, then the ghost fill */
int calc_u2p(void)
{
  state_in();
  if (koral_hip_calc_u2p(KG)) fail("calc_u2p");
  state_out();
  return 0;
}

/* finite.c:605-636: polar-axis correction */
int do_correct(void)
{
  state_in();
  if (koral_hip_do_correct(KG)) fail("do_correct");
  state_out();
  return 0;
}
