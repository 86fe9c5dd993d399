/*
 * Oracle: batched wrappers over independent points, called from the tests
 * (ctypes) and from the CPU-baseline leg of bench.py.  Test infrastructure
 * only.  Point i reads pp[i*nv..], uu[i*nv..], metric g20[i*20..], G20[i*20..].
 */
#include <stddef.h>
#include <omp.h>
#include "oracle.h"

int oracle_params_size(void) { return (int)sizeof(oracle_params); }

int oracle_params_offset(int field)
{
  switch (field) {
  case 0: return (int)offsetof(oracle_params, gamma);
  case 1: return (int)offsetof(oracle_params, kappa_model);
  case 2: return (int)offsetof(oracle_params, rhofloor);
  case 3: return (int)offsetof(oracle_params, u2pconv);
  case 4: return (int)offsetof(oracle_params, radimpmaxiter);
  case 5: return (int)offsetof(oracle_params, radimpconvrelentrerr);
  case 6: return (int)offsetof(oracle_params, int_order);
  case 7: return (int)offsetof(oracle_params, doradimpfixups);
  default: return -1;
  }
}

void oracle_p2u_batch(const oracle_params *P, long n, const double *pp, double *uu, const double *g20,
                      const double *G20)
{
  int nv = P->nv;
  for (long i = 0; i < n; i++) {
    o_geom g;
    o_fill_geom(&g, g20 + 20 * i, G20 + 20 * i);
    o_p2u(P, pp + nv * i, uu + nv * i, &g);
  }
}

/* calc_primitives (u2p.c:13-133) on independent points: u2p + floors + p2u;
   flags[i*7..] receives ENTROPYFLAG, ENTROPYFLAG2, HDFIXUPFLAG, RADFIXUPFLAG */
void oracle_u2p_batch(const oracle_params *P, long n, double *uu, double *pp, const double *g20,
                      const double *G20, int *ret, int *flags)
{
  int nv = P->nv;
  for (long i = 0; i < n; i++) {
    o_geom g;
    o_fill_geom(&g, g20 + 20 * i, G20 + 20 * i);
    double *u = uu + nv * i, *p = pp + nv * i;
    int *fl = flags + O_NFLAGS * i;
    int corrected[3] = {0, 0, 0}, fixups[2] = {0, 0};
    fl[O_ENTROPYFLAG] = 0;
    fl[O_ENTROPYFLAG2] = 0;
    ret[i] = o_u2p(P, u, p, &g, corrected, fixups);
    if (corrected[0] == 1) fl[O_ENTROPYFLAG] = 1;
    if (corrected[2] == 1) fl[O_ENTROPYFLAG2] = 1;
    o_check_floors_mhd(P, p, &g);
    if (P->rad) o_check_floors_rad(P, p, &g);
    o_p2u(P, p, u, &g);
    fl[O_HDFIXUPFLAG] = (fixups[0] > 0) ? 1 : 0;
    fl[O_RADFIXUPFLAG] = (fixups[1] > 0) ? -1 : 0;
  }
}

void oracle_u2p_solver_batch(const oracle_params *P, long n, const double *uu, double *pp,
                             const double *g20, const double *G20, int etype, int *ret)
{
  int nv = P->nv;
  for (long i = 0; i < n; i++) {
    o_geom g;
    o_fill_geom(&g, g20 + 20 * i, G20 + 20 * i);
    ret[i] = o_u2p_solver_W(P, uu + nv * i, pp + nv * i, &g, etype);
  }
}

/* The W inversion from a learned W start W0[i] (> 0; the plain start
   otherwise) under the unchanged tolerance; a start that fails is retried
   from the plain start, as the chain would run without it.  iters[i] counts
   the Newton iterations of both tries, guessed[i] is 1 when the learned start
   solved the point.  No reference counterpart. */
void oracle_u2p_solver_batch_from(const oracle_params *P, long n, const double *uu, double *pp,
                                  const double *g20, const double *G20, int etype, const double *W0, int *ret,
                                  int *iters, int *guessed)
{
  int nv = P->nv;
  for (long i = 0; i < n; i++) {
    o_geom g;
    o_fill_geom(&g, g20 + 20 * i, G20 + 20 * i);
    int it = 0, it2 = 0, r = -1, gd = 0;
    if (W0 && W0[i] > 0.) {
      r = o_u2p_solver_W_from(P, uu + nv * i, pp + nv * i, &g, etype, W0[i], &it);
      gd = r >= 0;
    }
    if (!gd) {
      r = o_u2p_solver_W_from(P, uu + nv * i, pp + nv * i, &g, etype, 0., &it2);
      it += it2;
    }
    ret[i] = r;
    if (iters) iters[i] = it;
    if (guessed) guessed[i] = gd;
  }
}

void oracle_u2p_rad_batch(const oracle_params *P, long n, const double *uu, double *pp,
                          const double *g20, const double *G20, int *ret, int *corr)
{
  int nv = P->nv;
  for (long i = 0; i < n; i++) {
    o_geom g;
    o_fill_geom(&g, g20 + 20 * i, G20 + 20 * i);
    ret[i] = o_u2p_rad(P, uu + nv * i, pp + nv * i, &g, corr + i);
  }
}

void oracle_floors_batch(const oracle_params *P, long n, double *pp, const double *g20, const double *G20,
                         int *ret)
{
  int nv = P->nv;
  for (long i = 0; i < n; i++) {
    o_geom g;
    o_fill_geom(&g, g20 + 20 * i, G20 + 20 * i);
    int r = o_check_floors_mhd(P, pp + nv * i, &g);
    int r2 = P->rad ? o_check_floors_rad(P, pp + nv * i, &g) : 0;
    ret[i] = (r < 0 ? 1 : 0) | (r2 < 0 ? 2 : 0);
  }
}

void oracle_Gi_batch(const oracle_params *P, long n, const double *pp, const double *g20, const double *G20,
                     int labframe, double *Gi)
{
  int nv = P->nv;
  for (long i = 0; i < n; i++) {
    o_geom g;
    o_fill_geom(&g, g20 + 20 * i, G20 + 20 * i);
    o_calc_Gi(P, pp + nv * i, &g, Gi + 4 * i, labframe);
  }
}

void oracle_ff_Rtt_batch(const oracle_params *P, long n, const double *pp, const double *g20,
                         const double *G20, double *Rtt)
{
  int nv = P->nv;
  for (long i = 0; i < n; i++) {
    o_geom g;
    o_fill_geom(&g, g20 + 20 * i, G20 + 20 * i);
    o_calc_ff_Rtt(P, pp + nv * i, &g, Rtt + i, NULL);
  }
}

/* implicit_lab_rad_source_term over independent cells; info[i*6..] =
   {iters, attempt, whichprim, whicheq, whichframe, total_iters} */
void oracle_implicit_batch_guess(const oracle_params *P, long n, double *uu, double *pp, const double *g20,
                                 const double *G20, const double *dt, double global_time, int *flags, int *info,
                                 int nthreads, const double *gxg, const double *gxr, int maxiter)
{
  int nv = P->nv;
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
  for (long i = 0; i < n; i++) {
    o_geom g;
    o_fill_geom(&g, g20 + 20 * i, G20 + 20 * i);
    o_imp_info inf;
    o_solve_implicit_lab_guess(P, uu + nv * i, pp + nv * i, &g, dt[i], global_time, flags + O_NFLAGS * i, &inf,
                               gxg ? gxg + 4 * i : NULL, gxr ? gxr + 4 * i : NULL, maxiter);
    int *o = info + 6 * i;
    o[0] = inf.iters;
    o[1] = inf.attempt;
    o[2] = inf.whichprim;
    o[3] = inf.whicheq;
    o[4] = inf.whichframe;
    o[5] = inf.total_iters;
  }
}

void oracle_implicit_batch(const oracle_params *P, long n, double *uu, double *pp, const double *g20,
                           const double *G20, const double *dt, double global_time, int *flags, int *info,
                           int nthreads)
{
  oracle_implicit_batch_guess(P, n, uu, pp, g20, G20, dt, global_time, flags, info, nthreads, NULL, NULL, 0);
}

/* ---- grid-level entry points (o_grid.c) ---- */
#include "o_grid.h"

int oracle_grid_set_bc(const oracle_params *P, const o_grid *G, double *u, double *p)
{
  o_set_bc(P, G, u, p);
  return 0;
}

double oracle_grid_tstepden(const oracle_params *P, const o_grid *G, const double *p) { return o_tstepden(P, G, p); }

int oracle_grid_explicit(const oracle_params *P, const o_grid *G, double t, double dt, double *u, double *p,
                         double *tstepdenmax)
{
  return o_op_explicit(P, G, t, dt, u, p, tstepdenmax);
}

int oracle_grid_explicit_parts(const oracle_params *P, const o_grid *G, double t, double dt, double *u, double *p,
                               double *tstepdenmax, double *ws, double *flx, double *fly, double *flz)
{
  double *fl[3] = {flx, fly, flz};
  return o_op_explicit_parts(P, G, t, dt, u, p, tstepdenmax, ws, fl);
}

int oracle_grid_cell_fixup(const oracle_params *P, const o_grid *G, double *u, double *p, int *flags, int type)
{
  return o_cell_fixup(P, G, u, p, flags, type);
}

int oracle_grid_u2p(const oracle_params *P, const o_grid *G, double *u, double *p, int *flags, int *counts)
{
  return o_calc_u2p(P, G, u, p, flags, counts);
}

int oracle_grid_implicit(const oracle_params *P, const o_grid *G, double t, double dt, double *u, double *p,
                         int *flags, int *counts)
{
  return o_op_implicit(P, G, t, dt, u, p, flags, counts);
}

int oracle_grid_implicit_guess(const oracle_params *P, const o_grid *G, double t, double dt, double *u, double *p,
                               int *flags, int *counts, const double *gxg, const double *gxr, int maxiter)
{
  return o_op_implicit_guess(P, G, t, dt, u, p, flags, counts, gxg, gxr, maxiter);
}

int oracle_grid_do_correct(const oracle_params *P, const o_grid *G, double *u, double *p)
{
  return o_do_correct(P, G, u, p);
}

int oracle_grid_step(const oracle_params *P, const o_grid *G, double t, double dt, double *u, double *p, int *flags,
                     double *tstepdenmax, int *counts)
{
  return o_step_rk2imex(P, G, t, dt, u, p, flags, tstepdenmax, counts);
}

/* threads of the grid operators' independent loops (1: serial) */
int oracle_grid_set_threads(int n)
{
  o_grid_set_threads(n);
  return 0;
}
