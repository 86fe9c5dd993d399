/*
 * koral_ops.h -- KORAL's operator entry points bound to libkoral_hip.so.
 *
 * This is the binding a KORAL maintainer adds to the reference (plain C,
 * built by KORAL's own gcc line): koral_ops.c implements KORAL's signatures
 *   int op_explicit(ldouble t, ldouble dt);   ko.h:455, finite.c:647
 *   int op_implicit(ldouble t, ldouble dt);   ko.h:456, finite.c:1242
 *   int calc_u2p(void);                       ko.h:444, finite.c:502
 *   int do_correct(void);                     ko.h:445, finite.c:605
 * on the device, in parity mode: each call copies KORAL's host arrays p, u
 * and cellflag in, runs the operator, and copies them back, so every caller
 * in ko.c stays unchanged.  In KORAL the globals below are its own (ko.h:
 * 116-143, 274-288); a build that includes ko.h drops these declarations.
 */
#ifndef KORAL_OPS_H
#define KORAL_OPS_H
#include "koral_hip.h"

typedef double ldouble;

/* KORAL's state: primitives and conserved [iz][iy][ix][NV] with NG ghosts,
   cell flags [iz][iy][ix][7], and the time-step density max of the step */
extern ldouble *p, *u;
extern int *cellflag;
extern ldouble tstepdenmax;

/* once after calc_metric: the resolved compile-time configuration and the
   tabulated metric; 0 on success */
int koral_ops_bind(const koral_params *P, const koral_grid_desc *D);
void koral_ops_unbind(void);
/* the device grid the operators run on (for the state set-up calls) */
koral_hip_grid *koral_ops_grid(void);

int op_explicit(ldouble t, ldouble dt);
int op_implicit(ldouble t, ldouble dt);
int calc_u2p(void);
int do_correct(void);

#endif
