/*
 * Oracle: grid-level operators of one KORAL timestep -- ghost fill, wave
 * speeds, reconstruction, fluxes, Riemann solver, flux-CT, conservative update
 * with metric sources, the u2p sweep with fixups, the implicit sweep, and the
 * RK2IMEX step body, restated on AoS arrays laid out as KORAL's
 * (ko.h:505-520): state[(iz*SY + iy)*SX + ix][iv] with ghosts.
 * Loops whose iterations each write only their own cell or face (wave speeds,
 * face sweeps, update, u2p, implicit, the stage algebra) may run on several
 * OpenMP threads (o_grid_threads, 1 by default, as KORAL's per-rank loops
 * run); results do not depend on the thread count.  The ghost fill and
 * cell_fixup, whose order matters, stay serial.
 * Test infrastructure only (see oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "o_grid.h"

#define SMALL 1.e-50
#define BIG (1. / SMALL)

/* threads of the independent loops (set by oracle_grid_set_threads) */
static int o_nthreads = 1;
void o_grid_set_threads(int n) { o_nthreads = n < 1 ? 1 : n; }

/* Work arrays of the operators, kept between calls and grown to the largest
   size asked for (KORAL keeps them as global arrays, ko.h:274-288; allocating
   them per call would time page faults).  Contents are not cleared: every
   entry an operator reads it has written in the same call.  Not re-entrant,
   as KORAL's globals are not. */
enum { W_PBL, W_PBR, W_FLL, W_FLR, W_WS, W_FLBX, W_FLBY, W_FLBZ, W_UT0, W_UT1, W_UT2, W_UFORGET, W_DUT1, W_DUT2,
       W_DRT1, W_DRT2, W_PN, W_UN, W_EMF, W_NSLOTS };
static double *o_wbuf[W_NSLOTS];
static size_t o_wcap[W_NSLOTS];
static double *o_work(int slot, size_t n)
{
  if (n > o_wcap[slot]) {
    free(o_wbuf[slot]);
    o_wbuf[slot] = malloc(sizeof(double) * n);
    o_wcap[slot] = o_wbuf[slot] ? n : 0;
  }
  return o_wbuf[slot];
}

static inline double mx(double a, double b) { return a > b ? a : b; }
static inline double mn(double a, double b) { return a < b ? a : b; } /* misc.c:56-60 */
static inline double sgn(double x) { return x > 0. ? 1. : (x < 0. ? -1. : 1.); }

/* ---- indexing (ko.h:225-245, 505-537) ---- */
static inline long cidx(const o_grid *G, int ix, int iy, int iz)
{
  long x = ix + G->ngx;
  long y = G->ny > 1 ? iy + G->ngy : 0;
  long z = G->nz > 1 ? iz + G->ngz : 0;
  return (z * G->sy + y) * G->sx + x;
}
static inline long fxidx(const o_grid *G, int ix, int iy, int iz)
{
  long y = G->ny > 1 ? iy + G->ngy : 0;
  long z = G->nz > 1 ? iz + G->ngz : 0;
  return (z * G->sy + y) * (G->sx + 1) + ix + G->ngx;
}
static inline long fyidx(const o_grid *G, int ix, int iy, int iz)
{
  long z = G->nz > 1 ? iz + G->ngz : 0;
  return (z * (G->sy + 1) + iy + G->ngy) * G->sx + ix + G->ngx;
}
static inline long fzidx(const o_grid *G, int ix, int iy, int iz)
{
  long y = G->ny > 1 ? iy + G->ngy : 0;
  return ((long)(iz + G->ngz) * G->sy + y) * G->sx + ix + G->ngx;
}
static inline long midx(const o_grid *G, int ix, int iy) { return (long)(G->ny > 1 ? iy + G->ngy : 0) * G->sx + ix + G->ngx; }

static double size_x(const o_grid *G, int ic, int d)
{
  const double *xb = G->xb[d];
  return xb[ic + 1 + G->ng] - xb[ic + G->ng];
}

void o_geom_cell(const o_grid *G, int ix, int iy, o_geom *g)
{
  long m = midx(G, ix, iy);
  o_fill_geom(g, G->gc + 20 * m, G->Gc + 20 * m);
}

void o_geom_face(const o_grid *G, int d, int ix, int iy, o_geom *g)
{
  if (d == 0) {
    long m = (long)(G->ny > 1 ? iy + G->ngy : 0) * (G->sx + 1) + ix + G->ngx;
    o_fill_geom(g, G->gfx + 20 * m, G->Gfx + 20 * m);
  } else if (d == 1) {
    long m = (long)(iy + G->ngy) * G->sx + ix + G->ngx;
    o_fill_geom(g, G->gfy + 20 * m, G->Gfy + 20 * m);
  } else {
    o_geom_cell(G, ix, iy, g);
  }
}

static int dim_n(const o_grid *G, int d) { return d == 0 ? G->nx : (d == 1 ? G->ny : G->nz); }

/* ---------------------------------------------------------------------------
 * Ghost fill: set_bc (finite.c:2752-3700).  Faces first (ghost cells with one
 * index outside), then the corners of the whole domain.
 * ------------------------------------------------------------------------- */
static void p2u_cell(const oracle_params *P, const o_grid *G, double *u, const double *p, int ix, int iy, int iz)
{
  o_geom g;
  o_geom_cell(G, ix, iy, &g);
  long c = cidx(G, ix, iy, iz);
  o_p2u(P, p + (long)P->nv * c, u + (long)P->nv * c, &g);
}

static void face_bc_cell(const oracle_params *P, const o_grid *G, double *u, double *p, int ix, int iy, int iz)
{
  int nv = P->nv;
  int ii[3] = {ix, iy, iz};
  int d = -1, hi = 0;
  for (int k = 0; k < 3; k++) {
    int n = dim_n(G, k);
    if (n > 1 && ii[k] < 0) { d = k; hi = 0; }
    if (n > 1 && ii[k] >= n) { d = k; hi = 1; }
  }
  int src[3] = {ix, iy, iz};
  int n = dim_n(G, d);
  int type = G->bc[2 * d + hi];
  int flip = 0;
  if (type == O_BC_PERIODIC) {
    if (src[d] < 0) src[d] += n;
    if (src[d] > n - 1) src[d] -= n;
  } else if (type == O_BC_REFLECT) {
    src[d] = hi ? 2 * n - src[d] - 1 : -src[d] - 1;
    flip = 1;
  } else {
    src[d] = hi ? n - 1 : 0;
  }
  double pv[O_NVMAX];
  const double *ps = p + (long)nv * cidx(G, src[0], src[1], src[2]);
  for (int iv = 0; iv < nv; iv++) pv[iv] = ps[iv];
  if (flip) {
    pv[2 + d] = -pv[2 + d];
    if (P->mag) pv[6 + d] = -pv[6 + d];
    if (P->rad) pv[6 + 3 * P->mag + 1 + d] = -pv[6 + 3 * P->mag + 1 + d];
  }
  long c = cidx(G, ix, iy, iz);
  for (int iv = 0; iv < nv; iv++) p[(long)nv * c + iv] = pv[iv];
  p2u_cell(P, G, u, p, ix, iy, iz);
}

static void copy_p(const oracle_params *P, const o_grid *G, double *p, int dx, int dy, int dz, int sx, int sy, int sz)
{
  int nv = P->nv;
  double *d = p + (long)nv * cidx(G, dx, dy, dz);
  const double *s = p + (long)nv * cidx(G, sx, sy, sz);
  for (int iv = 0; iv < nv; iv++) d[iv] = s[iv];
}

static void avg_p(const oracle_params *P, const o_grid *G, double *p, int dx, int dy, int dz, int ax, int ay,
                  int az, int bx, int by, int bz)
{
  int nv = P->nv;
  double *d = p + (long)nv * cidx(G, dx, dy, dz);
  const double *a = p + (long)nv * cidx(G, ax, ay, az);
  const double *b = p + (long)nv * cidx(G, bx, by, bz);
  for (int iv = 0; iv < nv; iv++) d[iv] = .5 * (a[iv] + b[iv]);
}

/* one edge of corners between dims a and b (finite.c:3336-3600): at every
   position along the third dim, copy the far ghosts of the edge rows from the
   first interior rows and fill the two diagonal cells by averaging */
static void corner_edge(const oracle_params *P, const o_grid *G, double *u, double *p, int a, int b, int ahi, int bhi)
{
  int NG = G->ng;
  int c = 3 - a - b;
  int na = dim_n(G, a), nb = dim_n(G, b), nc = dim_n(G, c);
  int ea = ahi ? na : -1, ia = ahi ? na - 1 : 0, ea2 = ahi ? na + 1 : -2;
  int eb = bhi ? nb : -1, ib = bhi ? nb - 1 : 0, eb2 = bhi ? nb + 1 : -2;
  for (int k = -NG; k < nc + NG; k++) {
    int q[3];
#define SET3(v, A, B, C) (v[a] = (A), v[b] = (B), v[c] = (C))
    for (int i = 0; i < NG - 1; i++) {
      int ga = ahi ? na + i + 1 : -NG + i;
      int gb = bhi ? nb + i + 1 : -NG + i;
      int d1[3], s1[3], d2[3], s2[3];
      SET3(d1, ga, eb, k); SET3(s1, ga, ib, k);
      SET3(d2, ea, gb, k); SET3(s2, ia, gb, k);
      copy_p(P, G, p, d1[0], d1[1], d1[2], s1[0], s1[1], s1[2]);
      copy_p(P, G, p, d2[0], d2[1], d2[2], s2[0], s2[1], s2[2]);
      p2u_cell(P, G, u, p, d1[0], d1[1], d1[2]);
      p2u_cell(P, G, u, p, d2[0], d2[1], d2[2]);
    }
    int dd[3], aa[3], bb[3];
    SET3(dd, ea, eb, k); SET3(aa, ea, ib, k); SET3(bb, ia, eb, k);
    avg_p(P, G, p, dd[0], dd[1], dd[2], aa[0], aa[1], aa[2], bb[0], bb[1], bb[2]);
    p2u_cell(P, G, u, p, dd[0], dd[1], dd[2]);
    SET3(dd, ea2, eb2, k); SET3(aa, ea2, eb, k); SET3(bb, ea, eb2, k);
    avg_p(P, G, p, dd[0], dd[1], dd[2], aa[0], aa[1], aa[2], bb[0], bb[1], bb[2]);
    p2u_cell(P, G, u, p, dd[0], dd[1], dd[2]);
    (void)q;
#undef SET3
  }
}

void o_set_bc(const oracle_params *P, const o_grid *G, double *u, double *p)
{
  int NG = G->ng;
  int x0 = -G->ngx, x1 = G->nx + G->ngx, y0 = -G->ngy, y1 = G->ny + G->ngy, z0 = -G->ngz, z1 = G->nz + G->ngz;
  /* faces: loop_2 = ghost cells that are outside in exactly one dimension
     (each reads interior cells only, so the order does not matter) */
#pragma omp parallel for collapse(3) schedule(dynamic, 256) num_threads(o_nthreads)
  for (int ix = x0; ix < x1; ix++)
    for (int iy = y0; iy < y1; iy++)
      for (int iz = z0; iz < z1; iz++) {
        int out = (ix < 0 || ix >= G->nx) + (G->ny > 1 && (iy < 0 || iy >= G->ny)) +
                  (G->nz > 1 && (iz < 0 || iz >= G->nz));
        if (out == 1) face_bc_cell(P, G, u, p, ix, iy, iz);
      }
  if (G->ny > 1) {
    /* corners along z (x-y edges), along y (x-z), along x (y-z) */
    for (int ah = 0; ah < 2; ah++)
      for (int bh = 0; bh < 2; bh++) corner_edge(P, G, u, p, 0, 1, ah, bh);
    if (G->nz > 1) {
      for (int ah = 0; ah < 2; ah++)
        for (int bh = 0; bh < 2; bh++) corner_edge(P, G, u, p, 0, 2, ah, bh);
      for (int ah = 0; ah < 2; ah++)
        for (int bh = 0; bh < 2; bh++) corner_edge(P, G, u, p, 1, 2, ah, bh);
      /* total corners, finite.c:3603-3700: copy of the domain corner cell */
      for (int hx = 0; hx < 2; hx++)
        for (int hy = 0; hy < 2; hy++)
          for (int hz = 0; hz < 2; hz++) {
            int sx = hx ? G->nx - 1 : 0, sy = hy ? G->ny - 1 : 0, sz = hz ? G->nz - 1 : 0;
            for (int ix = hx ? G->nx : -NG; ix < (hx ? G->nx + NG : 0); ix++)
              for (int iy = hy ? G->ny : -NG; iy < (hy ? G->ny + NG : 0); iy++)
                for (int iz = hz ? G->nz : -NG; iz < (hz ? G->nz + NG : 0); iz++) {
                  copy_p(P, G, p, ix, iy, iz, sx, sy, sz);
                  p2u_cell(P, G, u, p, ix, iy, iz);
                }
          }
    }
  }
}

/* ---------------------------------------------------------------------------
 * Reconstruction avg2point (finite.c:90-289) for one variable.
 * ------------------------------------------------------------------------- */
static double minmod3(double a, double b, double c) /* finite.c:30-49 */
{
  if (isnan(a) || isnan(b) || isnan(c)) return 1.;
  double x1;
  if (a < 0. && b < 0. && c < 0.) {
    x1 = (a > b) ? a : b;
    if (c > x1) x1 = c;
    return x1;
  } else if (a > 0. && b > 0. && c > 0.) {
    x1 = (a < b) ? a : b;
    if (c < x1) x1 = c;
    return x1;
  }
  return 0.;
}

static double DMM(double X, double Y) { return .5 * (sgn(X) + sgn(Y)) * mn(fabs(X), fabs(Y)); }
static double DM4(double W, double X, double Y, double Z)
{
  return 0.125 * (sgn(W) + sgn(X)) * fabs((sgn(W) + sgn(Y)) * (sgn(W) + sgn(Z))) *
         mn(mn(fabs(W), fabs(X)), mn(fabs(Y), fabs(Z)));
}

/* MP5 face value on the +side of cell 0 from (um2,um1,u0,up1,up2) */
static double mp5(double um2, double um1, double u0, double up1, double up2)
{
  const double alpha = 4., epsm = 1.e-10, BB2 = 1.3333333;
  double VOR = (3. * um2 - 20. * um1 + 90. * u0 + 60. * up1 - 5. * up2) / 128.;
  double VMP = u0 + DMM(up1 - u0, alpha * (u0 - um1));
  if ((VOR - u0) * (VOR - VMP) < epsm) return VOR;
  double DJM1 = um2 - 2. * um1 + u0;
  double DJ = um1 - 2. * u0 + up1;
  double DJP1 = u0 - 2. * up1 + up2;
  double DM4JPH = DM4(4. * DJ - DJP1, 4. * DJP1 - DJ, DJ, DJP1);
  double DM4JMH = DM4(4. * DJ - DJM1, 4. * DJM1 - DJ, DJ, DJM1);
  double VUL = u0 + alpha * (u0 - um1);
  double VAV = 0.5 * (u0 + up1);
  double VMD = VAV - 0.5 * DM4JPH;
  double VLC = u0 + 0.5 * (u0 - um1) + BB2 * DM4JMH;
  double VMIN = mx(mn(u0, mn(up1, VMD)), mn(u0, mn(VUL, VLC)));
  double VMAX = mn(mx(u0, mx(up1, VMD)), mx(u0, mx(VUL, VLC)));
  return VOR + DMM(VMIN - VOR, VMAX - VOR);
}

/* fills ul (left face) and ur (right face) of the centre cell */
static void avg2point(const oracle_params *P, const double *um2, const double *um1, double *u0, const double *up1,
                      const double *up2, double *ul, double *ur, double dxm2, double dxm1, double dx0, double dxp1,
                      double dxp2)
{
  int nv = P->nv;
  if (P->int_order == 1) {
    double th = P->minmod_theta;
    for (int i = 0; i < nv; i++) {
      ur[i] = u0[i] + .5 * minmod3(th * (up1[i] - u0[i]), .5 * (up1[i] - um1[i]), th * (u0[i] - um1[i]));
      ul[i] = u0[i] - .5 * minmod3(th * (up1[i] - u0[i]), .5 * (up1[i] - um1[i]), th * (u0[i] - um1[i]));
    }
  } else if (P->int_order == 2) {
    double dri[O_NVMAX], drim1[O_NVMAX], drip1[O_NVMAX];
    double Z1 = (dxm1 + dx0) / (2. * dx0 + dxp1);
    double Z2 = (dxp2 + dxp1) / (2. * dxp1 + dx0);
    double DX = dxm1 + dx0 + dxp1 + dxp2;
    for (int iv = 0; iv < nv; iv++) {
      dri[iv] = dx0 / (dxm1 + dx0 + dxp1) *
                ((2. * dxm1 + dx0) / (dxp1 + dx0) * (up1[iv] - u0[iv]) + (dx0 + 2. * dxp1) / (dxm1 + dx0) * (u0[iv] - um1[iv]));
      drip1[iv] = dxp1 / (dx0 + dxp1 + dxp2) *
                  ((2. * dx0 + dxp1) / (dxp2 + dxp1) * (up2[iv] - up1[iv]) + (dxp1 + 2. * dxp2) / (dx0 + dxp1) * (up1[iv] - u0[iv]));
      drim1[iv] = dxm1 / (dxm2 + dxm1 + dx0) *
                  ((2. * dxm2 + dxm1) / (dx0 + dxm1) * (u0[iv] - um1[iv]) + (dxm1 + 2. * dx0) / (dxm2 + dxm1) * (um1[iv] - um2[iv]));
      if ((up1[iv] - u0[iv]) * (u0[iv] - um1[iv]) > 0.)
        dri[iv] = mn(fabs(dri[iv]), mn(2. * fabs(u0[iv] - um1[iv]), 2. * fabs(u0[iv] - up1[iv]))) * sgn(dri[iv]);
      else
        dri[iv] = 0.;
      if ((up2[iv] - up1[iv]) * (up1[iv] - u0[iv]) > 0.)
        drip1[iv] = mn(fabs(drip1[iv]), mn(2. * fabs(up1[iv] - u0[iv]), 2. * fabs(up1[iv] - up2[iv]))) * sgn(drip1[iv]);
      else
        drip1[iv] = 0.;
      if ((u0[iv] - um1[iv]) * (um1[iv] - um2[iv]) > 0.)
        drim1[iv] = mn(fabs(drim1[iv]), mn(2. * fabs(um1[iv] - um2[iv]), 2. * fabs(um1[iv] - u0[iv]))) * sgn(drim1[iv]);
      else
        drim1[iv] = 0.;
      ur[iv] = u0[iv] + dx0 / (dx0 + dxp1) * (up1[iv] - u0[iv]) +
               1. / DX * ((2. * dxp1 * dx0) / (dxp1 + dx0) * (Z1 - Z2) * (up1[iv] - u0[iv]) - dx0 * Z1 * drip1[iv] + dxp1 * Z2 * dri[iv]);
    }
    Z1 = (dxm2 + dxm1) / (2. * dxm1 + dx0);
    Z2 = (dxp1 + dx0) / (2. * dx0 + dxm1);
    DX = dxm2 + dxm1 + dx0 + dxp1;
    for (int iv = 0; iv < nv; iv++)
      ul[iv] = um1[iv] + dxm1 / (dxm1 + dx0) * (u0[iv] - um1[iv]) +
               1. / DX * ((2. * dx0 * dxm1) / (dx0 + dxm1) * (Z1 - Z2) * (u0[iv] - um1[iv]) - dxm1 * Z1 * dri[iv] + dx0 * Z2 * drim1[iv]);
    for (int iv = 0; iv < nv; iv++) {
      if ((ur[iv] - u0[iv]) * (u0[iv] - ul[iv]) <= 0.) {
        ul[iv] = u0[iv];
        ur[iv] = u0[iv];
      }
      if ((ur[iv] - ul[iv]) * (ul[iv] - (3. * u0[iv] - 2. * ur[iv])) < 0.) ul[iv] = 3. * u0[iv] - 2. * ur[iv];
      if ((ur[iv] - ul[iv]) * ((3. * u0[iv] - 2. * ul[iv]) - ur[iv]) < 0.) ur[iv] = 3. * u0[iv] - 2. * ul[iv];
      u0[iv] = ul[iv] + .5 * (ur[iv] - ul[iv] + 6. * (u0[iv] - .5 * (ul[iv] + ur[iv])) * (1. - .5));
    }
  } else if (P->int_order == 4) {
    for (int iv = 0; iv < nv; iv++) {
      ur[iv] = mp5(um2[iv], um1[iv], u0[iv], up1[iv], up2[iv]);
      ul[iv] = mp5(up2[iv], up1[iv], u0[iv], um1[iv], um2[iv]);
    }
  } else {
    for (int i = 0; i < nv; i++) ur[i] = ul[i] = u0[i];
  }
}

/* ---------------------------------------------------------------------------
 * Fluxes f_flux_prime (physics.c:687-820) + f_flux_prime_rad (rad.c:3278-3330)
 * ------------------------------------------------------------------------- */
static void flux_prime(const oracle_params *P, const double *pp, int idim, const o_geom *g, double *ff)
{
  int nv = P->nv;
  for (int iv = 0; iv < nv; iv++) ff[iv] = 0.;
  double gdetu = g->gdet;
  double vr[4] = {0., pp[2], pp[3], pp[4]}, ucon[4], ucov[4];
  double bcon[4] = {0., 0., 0., 0.}, bcov[4] = {0., 0., 0., 0.}, bsq = 0.;
  o_relr_to_4vel(vr, ucon, ucov, g);
  if (P->mag) {
    o_bcon(pp, ucon, ucov, bcon, 6);
    o_lower(bcon, bcov, g->gg);
    bsq = bcon[0] * bcov[0] + bcon[1] * bcov[1] + bcon[2] * bcov[2] + bcon[3] * bcov[3];
  }
  double rho = pp[0], u = pp[1], S = pp[5];
  double pre = (P->gamma - 1.) * u;
  double w = rho + u + pre;
  double eta = w + bsq;
  double etap = u + pre + bsq;
  double ptot = pre + 0.5 * bsq;
  /* calc_Tij (physics.c:828-892) then indices_2221: T^i_j = T^{ik} g_kj */
  double T[4][4], Tm[4][4];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) T[i][j] = eta * ucon[i] * ucon[j] + ptot * g->GG[i][j] - bcon[i] * bcon[j];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      Tm[i][j] = 0.;
      for (int k = 0; k < 4; k++) Tm[i][j] += T[i][k] * g->gg[k][j];
    }
  double up1 = o_utp1(vr, g);
  int d = idim + 1;
  ff[0] = gdetu * rho * ucon[d];
  ff[1] = gdetu * (etap * ucon[d] * ucov[0] + rho * ucon[d] * up1);
  if (P->mag) ff[1] += -gdetu * bcon[d] * bcov[0];
  ff[2] = gdetu * (Tm[d][1]);
  ff[3] = gdetu * (Tm[d][2]);
  ff[4] = gdetu * (Tm[d][3]);
  ff[5] = gdetu * S * ucon[d];
  if (P->mag) {
    ff[6] = gdetu * (bcon[1] * ucon[d] - bcon[d] * ucon[1]);
    ff[7] = gdetu * (bcon[2] * ucon[d] - bcon[d] * ucon[2]);
    ff[8] = gdetu * (bcon[3] * ucon[d] - bcon[d] * ucon[3]);
  }
  if (P->rad) {
    int ee = 6 + 3 * P->mag;
    double R[4][4];
    o_calc_Rij(P, pp, g, R);
    double Rm[4];
    for (int j = 0; j < 4; j++) {
      Rm[j] = 0.;
      for (int k = 0; k < 4; k++) Rm[j] += R[d][k] * g->gg[k][j];
    }
    ff[ee] = gdetu * (Rm[0]);
    ff[ee + 1] = gdetu * (Rm[1]);
    ff[ee + 2] = gdetu * (Rm[2]);
    ff[ee + 3] = gdetu * (Rm[3]);
  }
}

/* ---------------------------------------------------------------------------
 * Wave speeds: calc_wavespeeds_lr_core/pure (physics.c:10-241) and
 * calc_rad_wavespeeds (rad.c:2451-2560)
 * ------------------------------------------------------------------------- */
/* returns without writing when the gas discriminant is negative: the
   reference then keeps whatever aret[] held (physics.c:43, 155-168) */
static void ws_core(const double *ucon, const o_geom *g, double w2, int idim, double *a0, double *a1, int rad)
{
  double Acov[4] = {0., 0., 0., 0.}, Acon[4], Bcov[4] = {1., 0., 0., 0.}, Bcon[4];
  Acov[idim + 1] = 1.;
  o_raise(Acov, Acon, g->GG);
  o_raise(Bcov, Bcon, g->GG);
  double Asq = Acon[0] * Acov[0] + Acon[1] * Acov[1] + Acon[2] * Acov[2] + Acon[3] * Acov[3];
  double Bsq = Bcon[0] * Bcov[0] + Bcon[1] * Bcov[1] + Bcon[2] * Bcov[2] + Bcon[3] * Bcov[3];
  double Au = Acov[0] * ucon[0] + Acov[1] * ucon[1] + Acov[2] * ucon[2] + Acov[3] * ucon[3];
  double Bu = Bcov[0] * ucon[0] + Bcov[1] * ucon[1] + Bcov[2] * ucon[2] + Bcov[3] * ucon[3];
  double AB = Acon[0] * Bcov[0] + Acon[1] * Bcov[1] + Acon[2] * Bcov[2] + Acon[3] * Bcov[3];
  double Au2 = Au * Au, Bu2 = Bu * Bu, AuBu = Au * Bu;
  double B = 2. * (AuBu * (1.0 - w2) - AB * w2);
  double A = Bu2 * (1.0 - w2) - Bsq * w2;
  double discr = 4.0 * w2 * ((AB * AB - Asq * Bsq) * w2 + (2.0 * AB * Au * Bu - Asq * Bu2 - Bsq * Au2) * (w2 - 1.0));
  if (discr < 0.) {
    if (!rad) return;
    discr = 0.;
  }
  discr = sqrt(discr);
  double c1 = -(-B + discr) / (2. * A);
  double c2 = -(-B - discr) / (2. * A);
  if (rad) {
    *a0 = mn(c1, c2);
    *a1 = mx(c1, c2);
  } else {
    *a0 = c1; /* physics.c:48-55 keeps the order in both branches */
    *a1 = c2;
  }
}

static void wavespeeds_cell(const oracle_params *P, const o_grid *G, const double *pp, int ix, int iy, int iz,
                            double *aaa)
{
  o_geom g;
  o_geom_cell(G, ix, iy, &g);
  double vr[4] = {0., pp[2], pp[3], pp[4]}, ucon[4], ucov[4];
  o_relr_to_4vel(vr, ucon, ucov, &g);
  double rho = pp[0], uu = pp[1];
  double pre = (P->gamma - 1.) * uu;
  double cs2 = P->gamma * pre / (rho + uu + pre);
  if (cs2 < 0.) cs2 = 0.;
  double va2 = 0.;
  if (P->mag) {
    double bcon[4], bcov[4];
    o_bcon(pp, ucon, ucov, bcon, 6);
    o_lower(bcon, bcov, g.gg);
    double bsq = bcon[0] * bcov[0] + bcon[1] * bcov[1] + bcon[2] * bcov[2] + bcon[3] * bcov[3];
    double EF = rho + P->gamma * uu;
    double EEloc = bsq + EF;
    va2 = bsq / EEloc;
    if (va2 < 0.) va2 = 0.;
  }
  double vtot2 = cs2 + va2 - cs2 * va2;
  double hd[6], rd[6] = {1., 1., 1., 1., 1., 1.}, aret[2] = {0., 0.};
  for (int d = 0; d < 3; d++) {
    ws_core(ucon, &g, vtot2, d, &aret[0], &aret[1], 0);
    hd[2 * d] = aret[0];
    hd[2 * d + 1] = aret[1];
  }
  if (P->rad) {
    int ee = 6 + 3 * P->mag;
    double dxp[3];
    int ii[3] = {ix, iy, iz};
    for (int d = 0; d < 3; d++) {
      double s = sqrt(g.gg[d + 1][d + 1]);
      dxp[d] = mx(size_x(G, ii[d], d) * s, size_x(G, ii[d] + 1, d) * s);
    }
    double chi = o_kappa_abs(P, pp) + o_kappa_es(P, pp);
    double ur[4] = {0., pp[ee + 1], pp[ee + 2], pp[ee + 3]}, urf[4];
    o_relr_to_4vel(ur, urf, NULL, &g);
    for (int d = 0; d < 3; d++) {
      double tau = chi * dxp[d];
      double rv2 = 1. / 3.;
      if (tau > 0.) {
        double rv2tau = 4. / 3. / tau * 4. / 3. / tau;
        rv2 = mn(rv2, rv2tau);
      }
      ws_core(urf, &g, rv2, d, &rd[2 * d], &rd[2 * d + 1], 1);
    }
  }
  for (int k = 0; k < 3; k++) {
    if (hd[2 * k] > 0.) hd[2 * k] = 0.;
    if (hd[2 * k + 1] < 0.) hd[2 * k + 1] = 0.;
    if (rd[2 * k] > 0.) rd[2 * k] = 0.;
    if (rd[2 * k + 1] < 0.) rd[2 * k + 1] = 0.;
  }
  for (int k = 0; k < 6; k++) {
    aaa[k] = hd[k];
    aaa[6 + k] = rd[k];
  }
}

/* per-cell wave-speed record: 0..5 hd l/r xyz, 6..11 rad l/r xyz, 12..14 hd
   max xyz, 15..17 rad max xyz (save_wavespeeds, finite.c:296-398) */
#define NWS 18

static double calc_wavespeeds(const oracle_params *P, const o_grid *G, const double *p, double *ws, double *tstepden)
{
  int nv = P->nv, lim = (P->int_order == 4) ? 2 : 1;
  int lx = lim, ly = G->ny > 1 ? lim : 0, lz = G->nz > 1 ? lim : 0;
  int corners = P->mag; /* MPI4CORNERS is forced by MAGNFIELD (choices.h:546-567) */
  double tmax = -1.;
#pragma omp parallel for collapse(3) schedule(dynamic, 64) reduction(max : tmax) num_threads(o_nthreads)
  for (int ix = -lx; ix < G->nx + lx; ix++)
    for (int iy = -ly; iy < G->ny + ly; iy++)
      for (int iz = -lz; iz < G->nz + lz; iz++) {
        int out = (ix < 0 || ix >= G->nx) + (iy < 0 || iy >= G->ny) + (iz < 0 || iz >= G->nz);
        if (!corners && out > 1) continue;
        long c = cidx(G, ix, iy, iz);
        double *a = ws + NWS * c;
        wavespeeds_cell(P, G, p + (long)nv * c, ix, iy, iz, a);
        for (int d = 0; d < 3; d++) {
          a[12 + d] = mx(fabs(a[2 * d]), fabs(a[2 * d + 1]));
          a[15 + d] = mx(fabs(a[6 + 2 * d]), fabs(a[6 + 2 * d + 1]));
        }
        if (out == 0) {
          double w[3];
          for (int d = 0; d < 3; d++) w[d] = P->rad ? mx(a[12 + d], a[15 + d]) : a[12 + d];
          double dx = size_x(G, ix, 0), dy = size_x(G, iy, 1), dz = size_x(G, iz, 2);
          double td;
          if (G->nz > 1 && G->ny > 1)
            td = w[0] / dx + w[1] / dy + w[2] / dz;
          else if (G->nz == 1 && G->ny > 1)
            td = w[0] / dx + w[1] / dy;
          else if (G->ny == 1 && G->nz > 1)
            td = w[0] / dx + w[2] / dz;
          else
            td = w[0] / dx;
          td /= P->tsteplim;
          if (tstepden) tstepden[c] = td;
          if (td > tmax) tmax = td;
        }
      }
  return tmax;
}

/* ---------------------------------------------------------------------------
 * op_explicit: finite.c:647-1234
 * ------------------------------------------------------------------------- */
static void face_pair_flux(const oracle_params *P, const o_grid *G, int d, const double *pL, const double *pR,
                           const double *fL, const double *fR, const double *wsL, const double *wsR, const o_geom *g,
                           double *out)
{
  /* f_calc_fluxes_at_faces, finite.c:1297-1415: am1 = left cell, ap1 = right cell */
  int nv = P->nv, nvmhd = 6 + 3 * P->mag;
  double uL[O_NVMAX], uR[O_NVMAX];
  o_p2u(P, pL, uL, g);
  o_p2u(P, pR, uR, g);
  for (int i = 0; i < nv; i++) {
    int r = (P->rad && i >= nvmhd) ? 1 : 0;
    double ag = mx(wsR[12 + 3 * r + d], wsL[12 + 3 * r + d]);
    double al = mn(wsR[6 * r + 2 * d], wsL[6 * r + 2 * d]);
    double ar = mx(wsR[6 * r + 2 * d + 1], wsL[6 * r + 2 * d + 1]);
    if (P->fluxmethod == 0) {
      out[i] = .5 * (fR[i] + fL[i] - ag * (uR[i] - uL[i]));
    } else {
      if (al > 0.)
        out[i] = fL[i];
      else if (ar < 0.)
        out[i] = fL[i];
      else
        out[i] = (-al * fR[i] + ar * fL[i] + al * ar * (uR[i] - uL[i])) / (ar - al);
    }
  }
}

static void sweep(const oracle_params *P, const o_grid *G, const double *p, const double *ws, int d, double *flb)
{
  /* reconstruction + face floors + one-sided fluxes for every cell of loop_1
     (finite.c:694-1014), then the Riemann solver per face */
  int nv = P->nv, NG = G->ng;
  int lim = (P->int_order == 4) ? 2 : 1;
  int n[3] = {G->nx, G->ny, G->nz};
  int l[3] = {lim, G->ny > 1 ? lim : 0, G->nz > 1 ? lim : 0};
  int t = P->mag ? 1 : 0; /* MPI4CORNERS: transverse rows -1..N */
  long nfaces = (d == 0) ? (long)(G->sx + 1) * G->sy * G->sz
                         : (d == 1 ? (long)G->sx * (G->sy + 1) * G->sz : (long)G->sx * G->sy * (G->sz + 1));
  double *pbL = o_work(W_PBL, nfaces * nv), *pbR = o_work(W_PBR, nfaces * nv);
  double *flL = o_work(W_FLL, nfaces * nv), *flR = o_work(W_FLR, nfaces * nv);
  (void)NG;
#pragma omp parallel for collapse(3) schedule(dynamic, 64) num_threads(o_nthreads)
  for (int ix = -l[0]; ix < n[0] + l[0]; ix++)
    for (int iy = -l[1]; iy < n[1] + l[1]; iy++)
      for (int iz = -l[2]; iz < n[2] + l[2]; iz++) {
        int ii[3] = {ix, iy, iz};
        int ok = 1;
        for (int k = 0; k < 3; k++) {
          if (k == d || n[k] == 1) continue;
          if (ii[k] < -t || ii[k] >= n[k] + t) ok = 0;
        }
        if (!P->mag) {
          int out = (ix < 0 || ix >= n[0]) + (iy < 0 || iy >= n[1]) + (iz < 0 || iz >= n[2]);
          if (out > 1) ok = 0;
        }
        if (!ok) continue;
        int dol = ii[d] >= 0, dor = ii[d] < n[d];
        const double *st[5];
        double dxs[5];
        for (int o = -2; o <= 2; o++) {
          int jj[3] = {ix, iy, iz};
          jj[d] += o;
          int need = (P->int_order > 1) || (o >= -1 && o <= 1);
          st[o + 2] = need ? p + (long)nv * cidx(G, jj[0], jj[1], jj[2]) : NULL;
          dxs[o + 2] = need ? size_x(G, jj[d], d) : 0.;
        }
        double u0[O_NVMAX], pl[O_NVMAX], pr[O_NVMAX], ffl[O_NVMAX], ffr[O_NVMAX];
        for (int iv = 0; iv < nv; iv++) u0[iv] = st[2][iv];
        const double *um2 = st[0] ? st[0] : st[1], *up2 = st[4] ? st[4] : st[3];
        avg2point(P, um2, st[1], u0, st[3], up2, pl, pr, dxs[0], dxs[1], dxs[2], dxs[3], dxs[4]);
        o_geom g;
        long fl = 0, fr = 0;
        int jr[3] = {ix, iy, iz};
        jr[d] += 1;
        if (d == 0) { fl = fxidx(G, ix, iy, iz); fr = fxidx(G, ix + 1, iy, iz); }
        if (d == 1) { fl = fyidx(G, ix, iy, iz); fr = fyidx(G, ix, iy + 1, iz); }
        if (d == 2) { fl = fzidx(G, ix, iy, iz); fr = fzidx(G, ix, iy, iz + 1); }
        if (dol) {
          o_geom_face(G, d, ix, iy, &g);
          o_check_floors_mhd(P, pl, &g);
          flux_prime(P, pl, d, &g, ffl);
        }
        if (dor) {
          o_geom_face(G, d, jr[0], jr[1], &g);
          o_check_floors_mhd(P, pr, &g);
          flux_prime(P, pr, d, &g, ffr);
        }
        for (int iv = 0; iv < nv; iv++) {
          pbR[fl * nv + iv] = pl[iv];
          pbL[fr * nv + iv] = pr[iv];
          if (dol) flR[fl * nv + iv] = ffl[iv];
          if (dor) flL[fr * nv + iv] = ffr[iv];
        }
      }
  /* Riemann solver on faces 0..N along d, rows -t..N+t-1 transverse */
  int lo[3], hi[3];
  for (int k = 0; k < 3; k++) {
    if (k == d) { lo[k] = 0; hi[k] = n[k] + 1; }
    else if (n[k] == 1) { lo[k] = 0; hi[k] = 1; }
    else { lo[k] = -t; hi[k] = n[k] + t; }
  }
#pragma omp parallel for collapse(3) schedule(dynamic, 64) num_threads(o_nthreads)
  for (int ix = lo[0]; ix < hi[0]; ix++)
    for (int iy = lo[1]; iy < hi[1]; iy++)
      for (int iz = lo[2]; iz < hi[2]; iz++) {
        long f = d == 0 ? fxidx(G, ix, iy, iz) : (d == 1 ? fyidx(G, ix, iy, iz) : fzidx(G, ix, iy, iz));
        int jm[3] = {ix, iy, iz};
        jm[d] -= 1;
        const double *wR = ws + NWS * cidx(G, ix, iy, iz), *wL = ws + NWS * cidx(G, jm[0], jm[1], jm[2]);
        o_geom g;
        o_geom_face(G, d, ix, iy, &g);
        face_pair_flux(P, G, d, pbL + f * nv, pbR + f * nv, flL + f * nv, flR + f * nv, wL, wR, &g, flb + f * nv);
      }
}

/* flux_ct: magn.c:199-339 (Toth EMF averaging) */
static void flux_ct(const oracle_params *P, const o_grid *G, double *flbx, double *flby, double *flbz)
{
  int nv = P->nv, NX = G->nx, NY = G->ny, NZ = G->nz;
  double c1 = (NY > 1 && NZ > 1) ? 0.25 : 0.5;
  double c2 = (NX > 1 && NZ > 1) ? 0.25 : 0.5;
  double c3 = (NX > 1 && NY > 1) ? 0.25 : 0.5;
  long ne = (long)(NX + 1) * (NY + 1) * (NZ + 1);
  double *emf = o_work(W_EMF, 3 * ne);
#define EMF(k, ix, iy, iz) emf[(k - 1) + 3 * ((ix) + (long)(iy) * (NX + 1) + (long)(iz) * (NY + 1) * (NX + 1))]
#define FX(iv, ix, iy, iz) flbx[fxidx(G, ix, NY > 1 ? iy : 0, NZ > 1 ? iz : 0) * nv + iv]
#define FY(iv, ix, iy, iz) flby[fyidx(G, NX > 1 ? ix : 0, iy, NZ > 1 ? iz : 0) * nv + iv]
#define FZ(iv, ix, iy, iz) flbz[fzidx(G, NX > 1 ? ix : 0, NY > 1 ? iy : 0, iz) * nv + iv]
  for (int ix = 0; ix <= NX; ix++)
    for (int iy = 0; iy <= NY; iy++)
      for (int iz = 0; iz <= NZ; iz++) {
        double e = 0.;
        if (NY > 1 || NZ > 1) {
          e = 0.;
          double s = 0.;
          if (NY > 1) s = s + FY(8, ix, iy, iz) + FY(8, ix, iy, iz - 1);
          if (NZ > 1) s = s - FZ(7, ix, iy, iz) - FZ(7, ix, iy - 1, iz);
          e = c1 * (s);
        }
        EMF(1, ix, iy, iz) = e;
        e = 0.;
        if (NX > 1 || NZ > 1) {
          double s = 0.;
          if (NZ > 1) s = s + FZ(6, ix, iy, iz) + FZ(6, ix - 1, iy, iz);
          if (NX > 1) s = s - FX(8, ix, iy, iz) - FX(8, ix, iy, iz - 1);
          e = c2 * (s);
        }
        EMF(2, ix, iy, iz) = e;
        e = 0.;
        if (NX > 1 || NY > 1) {
          double s = 0.;
          if (NX > 1) s = s + FX(7, ix, iy, iz) + FX(7, ix, iy - 1, iz);
          if (NY > 1) s = s - FY(6, ix, iy, iz) - FY(6, ix - 1, iy, iz);
          e = c3 * (s);
        }
        EMF(3, ix, iy, iz) = e;
      }
  for (int ix = 0; ix <= NX; ix++)
    for (int iy = 0; iy <= NY; iy++)
      for (int iz = 0; iz <= NZ; iz++) {
        if (NX > 1 && iy < NY && iz < NZ) {
          FX(6, ix, iy, iz) = 0.;
          FX(7, ix, iy, iz) = 0.5 * (EMF(3, ix, iy, iz) + EMF(3, ix, iy + 1, iz));
          FX(8, ix, iy, iz) = -0.5 * (EMF(2, ix, iy, iz) + EMF(2, ix, iy, iz + 1));
        }
        if (NY > 1 && ix < NX && iz < NZ) {
          FY(6, ix, iy, iz) = -0.5 * (EMF(3, ix, iy, iz) + EMF(3, ix + 1, iy, iz));
          FY(7, ix, iy, iz) = 0.;
          FY(8, ix, iy, iz) = 0.5 * (EMF(1, ix, iy, iz) + EMF(1, ix, iy, iz + 1));
        }
        if (NZ > 1 && ix < NX && iy < NY) {
          FZ(6, ix, iy, iz) = 0.5 * (EMF(2, ix, iy, iz) + EMF(2, ix + 1, iy, iz));
          FZ(7, ix, iy, iz) = -0.5 * (EMF(1, ix, iy, iz) + EMF(1, ix, iy + 1, iz));
          FZ(8, ix, iy, iz) = 0.;
        }
      }
#undef EMF
#undef FX
#undef FY
#undef FZ
}

/* f_metric_source_term_arb: physics.c:324-465 (GDETIN=1) */
static void metric_source(const oracle_params *P, const o_grid *G, const double *pp, int ix, int iy, double *ss)
{
  int nv = P->nv;
  o_geom g;
  o_geom_cell(G, ix, iy, &g);
  const double *K = G->kr + 64 * midx(G, ix, iy);
  double gdetu = g.gdet;
  double vr[4] = {0., pp[2], pp[3], pp[4]}, ucon[4], ucov[4];
  double bcon[4] = {0., 0., 0., 0.}, bsq = 0.;
  o_relr_to_4vel(vr, ucon, ucov, &g);
  if (P->mag) {
    double bcov[4];
    o_bcon(pp, ucon, ucov, bcon, 6);
    o_lower(bcon, bcov, g.gg);
    bsq = bcon[0] * bcov[0] + bcon[1] * bcov[1] + bcon[2] * bcov[2] + bcon[3] * bcov[3];
  }
  double pre = (P->gamma - 1.) * pp[1];
  double eta = pp[0] + pp[1] + pre + bsq;
  double ptot = pre + 0.5 * bsq;
  double T[4][4], Tm[4][4];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) T[i][j] = eta * ucon[i] * ucon[j] + ptot * g.GG[i][j] - bcon[i] * bcon[j];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      Tm[i][j] = 0.;
      for (int k = 0; k < 4; k++) Tm[i][j] += T[i][k] * g.gg[k][j];
    }
  double Rm[4][4];
  if (P->rad) {
    double R[4][4];
    o_calc_Rij(P, pp, &g, R);
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) {
        Rm[i][j] = 0.;
        for (int k = 0; k < 4; k++) Rm[i][j] += R[i][k] * g.gg[k][j];
      }
  }
  for (int iv = 0; iv < nv; iv++) ss[iv] = 0.;
  int ee = 6 + 3 * P->mag;
  for (int k = 0; k < 4; k++)
    for (int l = 0; l < 4; l++) {
      for (int j = 0; j < 4; j++) ss[1 + j] += gdetu * Tm[k][l] * K[l * 16 + j * 4 + k];
      if (P->rad)
        for (int j = 0; j < 4; j++) ss[ee + j] += gdetu * Rm[k][l] * K[l * 16 + j * 4 + k];
    }
}

int o_op_explicit(const oracle_params *P, const o_grid *G, double t, double dt, double *u, double *p,
                  double *tstepdenmax)
{
  return o_op_explicit_parts(P, G, t, dt, u, p, tstepdenmax, NULL, NULL);
}

/* op_explicit that also hands out its intermediate arrays: the wave speeds
   [cell][NWS] of the sweep and the face fluxes [face][nv] per direction
   after flux-CT (either pointer may be NULL). */
int o_op_explicit_parts(const oracle_params *P, const o_grid *G, double t, double dt, double *u, double *p,
                        double *tstepdenmax, double *ws_out, double **flb_out)
{
  int nv = P->nv;
  (void)t;
  o_set_bc(P, G, u, p);
  long ncell = (long)G->sx * G->sy * G->sz;
  double *ws = o_work(W_WS, ncell * NWS);
  /* the intermediates handed out are zero where the operator does not write */
  if (ws_out) memset(ws, 0, sizeof(double) * ncell * NWS);
  double tmax = calc_wavespeeds(P, G, p, ws, NULL);
  if (tstepdenmax && tmax > *tstepdenmax) *tstepdenmax = tmax;
  double *flb[3] = {NULL, NULL, NULL};
  long nf[3] = {(long)(G->sx + 1) * G->sy * G->sz, (long)G->sx * (G->sy + 1) * G->sz,
                (long)G->sx * G->sy * (G->sz + 1)};
  int n[3] = {G->nx, G->ny, G->nz};
  for (int d = 0; d < 3; d++) {
    if (n[d] == 1) continue;
    flb[d] = o_work(W_FLBX + d, nf[d] * nv);
    if (flb_out) memset(flb[d], 0, sizeof(double) * nf[d] * nv);
    sweep(P, G, p, ws, d, flb[d]);
  }
  if (P->mag) flux_ct(P, G, flb[0], flb[1], flb[2]);
#pragma omp parallel for collapse(3) schedule(dynamic, 64) num_threads(o_nthreads)
  for (int ix = 0; ix < G->nx; ix++)
    for (int iy = 0; iy < G->ny; iy++)
      for (int iz = 0; iz < G->nz; iz++) {
        long c = cidx(G, ix, iy, iz);
        double ms[O_NVMAX];
        metric_source(P, G, p + (long)nv * c, ix, iy, ms);
        double dx = size_x(G, ix, 0), dy = size_x(G, iy, 1), dz = size_x(G, iz, 2);
        for (int iv = 0; iv < nv; iv++) {
          double flxl = flb[0][fxidx(G, ix, iy, iz) * nv + iv], flxr = flb[0][fxidx(G, ix + 1, iy, iz) * nv + iv];
          double du = -(flxr - flxl) * dt / dx;
          if (G->ny > 1) {
            double fyl = flb[1][fyidx(G, ix, iy, iz) * nv + iv], fyr = flb[1][fyidx(G, ix, iy + 1, iz) * nv + iv];
            du = du - (fyr - fyl) * dt / dy;
          }
          if (G->nz > 1) {
            double fzl = flb[2][fzidx(G, ix, iy, iz) * nv + iv], fzr = flb[2][fzidx(G, ix, iy, iz + 1) * nv + iv];
            du = du - (fzr - fzl) * dt / dz;
          }
          u[(long)nv * c + iv] = u[(long)nv * c + iv] + du + ms[iv] * dt;
        }
      }
  if (ws_out) memcpy(ws_out, ws, sizeof(double) * ncell * NWS);
  if (flb_out)
    for (int d = 0; d < 3; d++)
      if (flb[d] && flb_out[d]) memcpy(flb_out[d], flb[d], sizeof(double) * nf[d] * nv);
  return 0;
}

/* cell_fixup: finite.c:4468-4640.  Loop order is ix, iy, iz (alloc_loops
   finite.c:1866-1878) and a processed cell's flag is cleared before later
   cells look at it, so lower neighbours always count as good. */
static void cell_fixup(const oracle_params *P, const o_grid *G, double *u, double *p, int *flags, int type)
{
  if (!P->dofixups) return;
  if (type == O_HDFIXUPFLAG && !P->dou2pmhdfixups) return;
  if (type == O_RADFIXUPFLAG && !P->dou2pradfixups) return;
  if (type == O_RADIMPFIXUPFLAG && !P->doradimpfixups) return;
  int nv = P->nv, nvmhd = 6 + 3 * P->mag, ee = nvmhd;
  long ncell = (long)G->sx * G->sy * G->sz;
  double *pn = o_work(W_PN, nv * ncell), *un = o_work(W_UN, nv * ncell);
  memcpy(pn, p, sizeof(double) * nv * ncell);
  memcpy(un, u, sizeof(double) * nv * ncell);
  int NX = G->nx, NY = G->ny, NZ = G->nz;
  for (int ix = 0; ix < NX; ix++)
    for (int iy = 0; iy < NY; iy++)
      for (int iz = 0; iz < NZ; iz++) {
        long c = cidx(G, ix, iy, iz);
        if (flags[O_NFLAGS * c + type] == 0) continue;
        flags[O_NFLAGS * c + type] = 0;
        const int nb[6][3] = {{ix - 1, iy, iz}, {ix + 1, iy, iz}, {ix, iy - 1, iz},
                              {ix, iy + 1, iz}, {ix, iy, iz - 1}, {ix, iy, iz + 1}};
        const int inb[6] = {ix - 1 >= 0, ix + 1 < NX, iy - 1 >= 0, iy + 1 < NY, iz - 1 >= 0, iz + 1 < NZ};
        const double *ppn[6];
        int in = 0;
        for (int k = 0; k < 6; k++) {
          if (!inb[k]) continue;
          long cn = cidx(G, nb[k][0], nb[k][1], nb[k][2]);
          if (flags[O_NFLAGS * cn + type] == 0) ppn[in++] = p + (long)nv * cn;
        }
        if ((NZ == 1 && NY == 1 && in >= 1) || (NZ == 1 && in >= 2) || (NY == 1 && in >= 2) || in >= 3) {
          double pp[O_NVMAX], uu[O_NVMAX];
          for (int iv = 0; iv < nv; iv++) {
            /* B1 = NVHD = 6 whether or not MAGNFIELD is on (mnemonics.h:23) */
            int fix = 0;
            if (type == O_HDFIXUPFLAG && iv != 0 && iv < 6) fix = 1;
            if (type == O_RADFIXUPFLAG && iv != 0 && P->rad && iv >= ee && iv <= ee + 3) fix = 1;
            if (type == O_RADIMPFIXUPFLAG && iv != 0 && (iv < 6 || (P->rad && iv >= ee && iv <= ee + 3))) fix = 1;
            if (fix) {
              pp[iv] = 0.;
              for (int q = 0; q < in; q++) pp[iv] += ppn[q][iv];
              pp[iv] /= (double)in;
            } else
              pp[iv] = p[(long)nv * c + iv];
          }
          o_geom g;
          o_geom_cell(G, ix, iy, &g);
          o_p2u(P, pp, uu, &g);
          for (int iv = 0; iv < nvmhd; iv++) {
            un[(long)nv * c + iv] = uu[iv];
            pn[(long)nv * c + iv] = pp[iv];
          }
        }
      }
  /* copyi_u back over the domain */
  for (int ix = 0; ix < NX; ix++)
    for (int iy = 0; iy < NY; iy++)
      for (int iz = 0; iz < NZ; iz++) {
        long c = cidx(G, ix, iy, iz);
        for (int iv = 0; iv < nv; iv++) {
          p[(long)nv * c + iv] = pn[(long)nv * c + iv];
          u[(long)nv * c + iv] = un[(long)nv * c + iv];
        }
      }

}

/* cell_fixup on its own (test entry of the oracle) */
int o_cell_fixup(const oracle_params *P, const o_grid *G, double *u, double *p, int *flags, int type)
{
  cell_fixup(P, G, u, p, flags, type);
  return 0;
}

/* calc_u2p: finite.c:502-561 with calc_primitives u2p.c:13-133 */
int o_calc_u2p(const oracle_params *P, const o_grid *G, double *u, double *p, int *flags, int *counts)
{
  int nv = P->nv;
  int c0 = 0, c1 = 0, c2 = 0;
#pragma omp parallel for collapse(3) schedule(dynamic, 64) reduction(+ : c0, c1, c2) num_threads(o_nthreads)
  for (int ix = 0; ix < G->nx; ix++)
    for (int iy = 0; iy < G->ny; iy++)
      for (int iz = 0; iz < G->nz; iz++) {
        long c = cidx(G, ix, iy, iz);
        int *fl = flags + O_NFLAGS * c;
        double *uu = u + (long)nv * c, *pp = p + (long)nv * c;
        o_geom g;
        o_geom_cell(G, ix, iy, &g);
        int corrected[3] = {0, 0, 0}, fixups[2] = {0, 0};
        fl[O_ENTROPYFLAG] = 0;
        fl[O_ENTROPYFLAG2] = 0;
        o_u2p(P, uu, pp, &g, corrected, fixups);
        if (corrected[0] == 1) fl[O_ENTROPYFLAG] = 1;
        if (corrected[2] == 1) fl[O_ENTROPYFLAG2] = 1;
        o_check_floors_mhd(P, pp, &g);
        if (P->rad) o_check_floors_rad(P, pp, &g);
        o_p2u(P, pp, uu, &g);
        fl[O_HDFIXUPFLAG] = fixups[0] > 0 ? 1 : 0;
        fl[O_RADFIXUPFLAG] = fixups[1] > 0 ? -1 : 0;
        c0 += fixups[0] > 0;
        c1 += fixups[1] > 0;
        c2 += corrected[0] == 1;
      }
  if (counts) {
    counts[0] += c0;
    counts[1] += c1;
    counts[2] += c2;
  }
  cell_fixup(P, G, u, p, flags, O_HDFIXUPFLAG);
  if (P->rad) cell_fixup(P, G, u, p, flags, O_RADFIXUPFLAG);
  o_set_bc(P, G, u, p);
  return 0;
}

/* op_implicit: finite.c:1242-1292 */
int o_op_implicit(const oracle_params *P, const o_grid *G, double t, double dt, double *u, double *p, int *flags,
                  int *counts)
{
  return o_op_implicit_guess(P, G, t, dt, u, p, flags, counts, NULL, NULL, 0);
}

/* op_implicit with learned Newton starts (no reference counterpart): gxg,
   gxr [cell][4] over the interior cells in x-fastest order, each solved by
   o_solve_implicit_lab_guess with the iteration budget maxiter; NULL gxg is
   op_implicit itself. */
int o_op_implicit_guess(const oracle_params *P, const o_grid *G, double t, double dt, double *u, double *p,
                        int *flags, int *counts, const double *gxg, const double *gxr, int maxiter)
{
  if (!P->rad) return 0;
  int nv = P->nv;
  int c0 = 0, c1 = 0, c2 = 0;
#pragma omp parallel for collapse(3) schedule(dynamic, 16) reduction(+ : c0, c1, c2) num_threads(o_nthreads)
  for (int ix = 0; ix < G->nx; ix++)
    for (int iy = 0; iy < G->ny; iy++)
      for (int iz = 0; iz < G->nz; iz++) {
        long c = cidx(G, ix, iy, iz);
        o_geom g;
        o_geom_cell(G, ix, iy, &g);
        o_imp_info inf;
        const long i = ix + (long)G->nx * (iy + (long)G->ny * iz);
        int r = o_solve_implicit_lab_guess(P, u + (long)nv * c, p + (long)nv * c, &g, dt, t, flags + O_NFLAGS * c,
                                           &inf, gxg ? gxg + 4 * i : NULL, gxr ? gxr + 4 * i : NULL, maxiter);
        c0 += (r == 0);
        c1 += inf.iters;
        c2 += (r != 0);
      }
  if (counts) {
    counts[0] += c0;
    counts[1] += c1;
    counts[2] += c2;
  }
  cell_fixup(P, G, u, p, flags, O_RADIMPFIXUPFLAG);
  return 0;
}

/* ---------------------------------------------------------------------------
 * do_correct (finite.c:605-636) with CORRECT_POLARAXIS: correct_polaraxis
 * (finite.c:4924-5060) for spherical-like coordinates.  In every (ix, iz)
 * column the nc cells next to each polar face take the gas and radiation
 * densities and the r/phi velocities of cell nc (NY-1-nc at the lower face),
 * the theta velocities scaled by |x2 - x2_axis| / |x2_src - x2_axis| (code
 * coordinates of cell centres, set_grid finite.c:1630-1660), then u = p2u(p)
 * for every variable except the field components.  The upper face is
 * processed before the lower one in each column, as the reference loops.
 * ------------------------------------------------------------------------- */
static double x_centre(const o_grid *G, int i, int d)
{
  const double *xb = G->xb[d];
  return .5 * (xb[i + 1 + G->ng] + xb[i + G->ng]);
}

int o_do_correct(const oracle_params *P, const o_grid *G, double *u, double *p)
{
  const int nc = P->polar_nc, nv = P->nv;
  if (nc <= 0 || G->ny < 2) return 0;
  const int ee = 6 + 3 * (P->mag ? 1 : 0);
  double pp[O_NVMAX], uu[O_NVMAX];
  for (int ix = 0; ix < G->nx; ix++)
    for (int iz = 0; iz < G->nz; iz++)
      for (int side = 0; side < 2; side++) {
        const double thaxis = side == 0 ? G->xb[1][0 + G->ng] : G->xb[1][G->ny + G->ng];
        const int iysrc = side == 0 ? nc : G->ny - 1 - nc;
        for (int ic = 0; ic < nc; ic++) {
          const int iy = side == 0 ? ic : G->ny - 1 - ic;
          const double th = x_centre(G, iy, 1), thsrc = x_centre(G, iysrc, 1);
          const double fac = fabs((th - thaxis) / (thsrc - thaxis));
          const double *ps = p + (long)nv * cidx(G, ix, iysrc, iz);
          double *pd = p + (long)nv * cidx(G, ix, iy, iz);
          double *ud = u + (long)nv * cidx(G, ix, iy, iz);
          for (int iv = 0; iv < nv; iv++) pp[iv] = pd[iv];
          pp[0] = ps[0];
          pp[1] = ps[1];
          pp[5] = ps[5];
          pp[2] = ps[2];
          pp[4] = ps[4];
          pp[3] = fac * ps[3];
          if (P->rad) {
            pp[ee] = ps[ee];
            pp[ee + 1] = ps[ee + 1];
            pp[ee + 3] = ps[ee + 3];
            pp[ee + 2] = fac * ps[ee + 2];
          }
          o_geom g;
          o_geom_cell(G, ix, iy, &g);
          o_p2u(P, pp, uu, &g);
          for (int iv = 0; iv < nv; iv++) {
            if (P->mag && iv >= 6 && iv <= 8) continue;
            pd[iv] = pp[iv];
            ud[iv] = uu[iv];
          }
        }
      }
  return 0;
}

/* RK2IMEX step body: ko.c:444-570 */
int o_step_rk2imex(const oracle_params *P, const o_grid *G, double t, double dt, double *u, double *p, int *flags,
                   double *tstepdenmax, int *counts)
{
  int nv = P->nv;
  long ncell = (long)G->sx * G->sy * G->sz;
  size_t bytes = sizeof(double) * nv * ncell;
  double gamma = 1. - 1. / sqrt(2.);
  const size_t nw = bytes / sizeof(double);
  double *ut0 = o_work(W_UT0, nw), *ut1 = o_work(W_UT1, nw), *ut2 = o_work(W_UT2, nw);
  double *uforget = o_work(W_UFORGET, nw), *dut1 = o_work(W_DUT1, nw), *dut2 = o_work(W_DUT2, nw);
  double *drt1 = o_work(W_DRT1, nw), *drt2 = o_work(W_DRT2, nw);
  double tmax = -1.;
#define DOM                                                                \
  _Pragma("omp parallel for collapse(3) schedule(static) num_threads(o_nthreads)") \
  for (int ix = 0; ix < G->nx; ix++) for (int iy = 0; iy < G->ny; iy++) for (int iz = 0; iz < G->nz; iz++)
#define C(a) a[(long)nv * cidx(G, ix, iy, iz) + iv]
  DOM for (int iv = 0; iv < nv; iv++) C(ut0) = 1. * C(u);
  o_op_implicit(P, G, t, dt * gamma, u, p, flags, counts);
  DOM for (int iv = 0; iv < nv; iv++) C(drt1) = (1. / (dt * gamma)) * C(u) + (-1. / (dt * gamma)) * C(ut0);
  DOM for (int iv = 0; iv < nv; iv++) C(ut1) = 1. * C(u);
  o_calc_u2p(P, G, u, p, flags, counts ? counts + 3 : NULL);
  o_do_correct(P, G, u, p);
  o_op_explicit(P, G, t, dt, u, p, &tmax);
  DOM for (int iv = 0; iv < nv; iv++) C(dut1) = (1. / (dt)) * C(u) + (-1. / (dt)) * C(ut1);
  DOM for (int iv = 0; iv < nv; iv++) C(u) = C(ut0) + (dt) * C(dut1) + (dt * (1. - 2. * gamma)) * C(drt1);
  DOM for (int iv = 0; iv < nv; iv++) C(uforget) = 1. * C(u);
  o_calc_u2p(P, G, u, p, flags, counts ? counts + 3 : NULL);
  o_do_correct(P, G, u, p);
  o_op_implicit(P, G, t, gamma * dt, u, p, flags, counts);
  DOM for (int iv = 0; iv < nv; iv++) C(drt2) = (1. / (dt * gamma)) * C(u) + (-1. / (dt * gamma)) * C(uforget);
  DOM for (int iv = 0; iv < nv; iv++) C(ut2) = 1. * C(u);
  o_calc_u2p(P, G, u, p, flags, counts ? counts + 3 : NULL);
  o_do_correct(P, G, u, p);
  o_op_explicit(P, G, t, dt, u, p, &tmax);
  DOM for (int iv = 0; iv < nv; iv++) C(dut2) = (1. / (dt)) * C(u) + (-1. / (dt)) * C(ut2);
  DOM for (int iv = 0; iv < nv; iv++) C(u) = C(ut0) + (dt / 2.) * C(dut1) + (dt / 2.) * C(dut2);
  DOM for (int iv = 0; iv < nv; iv++) C(u) = C(u) + (dt / 2.) * C(drt1) + (dt / 2.) * C(drt2);
  o_calc_u2p(P, G, u, p, flags, counts ? counts + 3 : NULL);
  o_do_correct(P, G, u, p);
#undef DOM
#undef C
  if (tstepdenmax) *tstepdenmax = tmax;
  return 0;
}

/* wave speeds + time-step density alone (used to seed dt before the first
   step, as ko.c does after the initial set_bc) */
double o_tstepden(const oracle_params *P, const o_grid *G, const double *p)
{
  long ncell = (long)G->sx * G->sy * G->sz;
  double *ws = calloc(ncell * NWS, sizeof(double));
  double t = calc_wavespeeds(P, G, p, ws, NULL);
  free(ws);
  return t;
}
