/* Oracle grid description (test infrastructure only). */
#ifndef KORAL_ORACLE_GRID_H
#define KORAL_ORACLE_GRID_H
#include "oracle.h"

#define O_BC_COPY 0
#define O_BC_PERIODIC 1
#define O_BC_REFLECT 2

/* Interior sizes nx,ny,nz (1 for a collapsed dimension), NG ghosts; a
   collapsed dimension carries no ghosts (ko.h:225-245).  Metric tables are
   per (x1,x2) point (no x3 dependence), AoS: gc[(iy*SX+ix)*20+c] etc.
   xb[d][i+ng] is the lower face of cell i (i from -ng to n+ng). */
typedef struct {
  int nx, ny, nz, ng, ngx, ngy, ngz, sx, sy, sz;
  int bc[6]; /* xlo, xhi, ylo, yhi, zlo, zhi */
  const double *xb[3];
  const double *gc, *Gc, *kr, *gfx, *Gfx, *gfy, *Gfy;
} o_grid;

void o_geom_cell(const o_grid *G, int ix, int iy, o_geom *g);
void o_geom_face(const o_grid *G, int d, int ix, int iy, o_geom *g);
void o_set_bc(const oracle_params *P, const o_grid *G, double *u, double *p);
int o_op_explicit(const oracle_params *P, const o_grid *G, double t, double dt, double *u, double *p,
                  double *tstepdenmax);
int o_op_explicit_parts(const oracle_params *P, const o_grid *G, double t, double dt, double *u, double *p,
                        double *tstepdenmax, double *ws_out, double **flb_out);
int o_cell_fixup(const oracle_params *P, const o_grid *G, double *u, double *p, int *flags, int type);
int o_calc_u2p(const oracle_params *P, const o_grid *G, double *u, double *p, int *flags, int *counts);
int o_op_implicit_guess(const oracle_params *P, const o_grid *G, double t, double dt, double *u, double *p,
                        int *flags, int *counts, const double *gxg, const double *gxr, int maxiter);
int o_op_implicit(const oracle_params *P, const o_grid *G, double t, double dt, double *u, double *p, int *flags,
                  int *counts);
int o_step_rk2imex(const oracle_params *P, const o_grid *G, double t, double dt, double *u, double *p, int *flags,
                   double *tstepdenmax, int *counts);
double o_tstepden(const oracle_params *P, const o_grid *G, const double *p);
int o_do_correct(const oracle_params *P, const o_grid *G, double *u, double *p);
void o_grid_set_threads(int n);
#endif
