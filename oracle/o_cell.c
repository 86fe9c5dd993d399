/*
 * Oracle: per-point geometry algebra, primitive->conserved, MHD inversion and
 * floors.  Test infrastructure only (see oracle.h).
 */
#include <math.h>
#include <float.h>
#include <stddef.h>
#include "oracle.h"
#include "o_math.h"

#define B1(P) 6
#define EE(P) (6 + 3 * (P)->mag)
#define SMALL 1.e-50
#define BIG (1. / SMALL)

static inline double mx(double a, double b) { return a > b ? a : b; }

/* fill_geometry: metric.c:1643-1676 */
void o_fill_geom(o_geom *g, const double *g20, const double *G20)
{
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 5; j++) {
      g->gg[i][j] = g20[i * 5 + j];
      g->GG[i][j] = G20[i * 5 + j];
    }
  g->alpha = sqrt(-1. / g->GG[0][0]);
  g->gdet = g->gg[3][4];
  g->gttpert = g->GG[3][4];
}

/* indices_21: frames.c:2086-2106 (A^i -> A_i) */
void o_lower(const double *a, double *b, const double gg[4][5])
{
  double t[4];
  for (int i = 0; i < 4; i++) {
    t[i] = 0.;
    for (int k = 0; k < 4; k++) t[i] += a[k] * gg[i][k];
  }
  for (int i = 0; i < 4; i++) b[i] = t[i];
}

/* indices_12: frames.c:2060-2080 (A_i -> A^i) */
void o_raise(const double *a, double *b, const double GG[4][5])
{
  double t[4];
  for (int i = 0; i < 4; i++) {
    t[i] = 0.;
    for (int k = 0; k < 4; k++) t[i] += a[k] * GG[i][k];
  }
  for (int i = 0; i < 4; i++) b[i] = t[i];
}

static double qsq_of(const double *v, const o_geom *g)
{
  double q = 0.;
  for (int i = 1; i < 4; i++)
    for (int j = 1; j < 4; j++) q += v[i] * v[j] * g->gg[i][j];
  return q;
}

/* conv_vels_both VELR -> VEL4: relele.c:105-130 with conv_vels_core
   relele.c:322-338 */
void o_relr_to_4vel(const double *ur, double *ucon, double *ucov, const o_geom *g)
{
  double qsq = qsq_of(ur, g);
  double gamma2 = 1. + qsq;
  double alpha2 = -1. / g->GG[0][0];
  double alpgam = sqrt(alpha2 * gamma2);
  double uc[4];
  uc[0] = sqrt(gamma2 / alpha2);
  for (int i = 1; i < 4; i++) uc[i] = ur[i] - alpgam * g->GG[0][i];
  double uv[4];
  o_lower(uc, uv, g->gg);
  for (int i = 0; i < 4; i++) {
    ucon[i] = uc[i];
    if (ucov) ucov[i] = uv[i];
  }
}

/* calc_utp1 (VELR branch): p2u.c:548-575 -- 1+u_t without cancellation */
double o_utp1(const double *vr, const o_geom *g)
{
  double qsq = qsq_of(vr, g);
  double gamma2 = 1. + qsq;
  double alphasq = -1. / g->GG[0][0];
  double alpgam = sqrt(alphasq * gamma2);
  double bsqoasq = g->gg[0][1] * g->GG[0][1] + g->gg[0][2] * g->GG[0][2] + g->gg[0][3] * g->GG[0][3];
  double ud0 = 0.0;
  for (int j = 1; j < 4; j++) ud0 += vr[j] * g->gg[0][j];
  return ud0 + (g->gttpert - alphasq * (bsqoasq + qsq)) / (1.0 + alpgam);
}

/* calc_bcon_4vel: magn.c:7-23 */
void o_bcon(const double *pp, const double *ucon, const double *ucov, double *bcon, int b1)
{
  bcon[0] = pp[b1] * ucov[1] + pp[b1 + 1] * ucov[2] + pp[b1 + 2] * ucov[3];
  for (int j = 1; j < 4; j++) bcon[j] = (pp[b1 - 1 + j] + bcon[0] * ucon[j]) / ucon[0];
}

static inline double dot4(const double *a, const double *b)
{
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}

/* calc_sfromu / calc_Sfromu: physics.c:907-917 */
double o_calc_Sfromu(const oracle_params *P, double rho, double u)
{
  double gm1 = P->gamma - 1.;
  double n = 1.0 / gm1;
  return rho * O_LOG(O_POW(gm1 * u, n) / O_POW(rho, n + 1.0));
}

/* p2u_mhd: p2u.c:56-159 */
int o_p2u_mhd(const oracle_params *P, const double *pp, double *uu, const o_geom *g)
{
  double gdetu = g->gdet;
  double rho = pp[0], u = pp[1];
  double vcon[4] = {0., pp[2], pp[3], pp[4]};
  double ucon[4], ucov[4], bcon[4] = {0., 0., 0., 0.}, bcov[4] = {0., 0., 0., 0.}, bsq = 0.;
  double S = pp[5];
  o_relr_to_4vel(vcon, ucon, ucov, g);
  if (P->mag) {
    o_bcon(pp, ucon, ucov, bcon, B1(P));
    o_lower(bcon, bcov, g->gg);
    bsq = dot4(bcon, bcov);
  }
  double ut = ucon[0];
  double rhout = rho * ut;
  double Sut = S * ut;
  double pre = (P->gamma - 1.) * u;
  double w = rho + u + pre;
  double eta = w + bsq;
  double etap = u + pre + bsq;
  double ptot = pre + 0.5 * bsq;
  double utp1 = o_utp1(vcon, g);
  double Tttt = etap * ucon[0] * ucov[0] + rho * ucon[0] * utp1 + ptot - bcon[0] * bcov[0];
  double Ttr = eta * ucon[0] * ucov[1] - bcon[0] * bcov[1];
  double Ttth = eta * ucon[0] * ucov[2] - bcon[0] * bcov[2];
  double Ttph = eta * ucon[0] * ucov[3] - bcon[0] * bcov[3];
  uu[0] = gdetu * rhout;
  uu[1] = gdetu * Tttt;
  uu[2] = gdetu * Ttr;
  uu[3] = gdetu * Ttth;
  uu[4] = gdetu * Ttph;
  uu[5] = gdetu * Sut;
  if (P->mag) {
    uu[6] = gdetu * pp[6];
    uu[7] = gdetu * pp[7];
    uu[8] = gdetu * pp[8];
  }
  return 0;
}

/* p2u_rad: p2u.c:253-307 (M1) */
int o_p2u_rad(const oracle_params *P, const double *pp, double *uu, const o_geom *g)
{
  int ee = EE(P);
  double gdetu = g->gdet;
  double Erf = pp[ee];
  double urf[4] = {0., pp[ee + 1], pp[ee + 2], pp[ee + 3]};
  double uc[4];
  o_relr_to_4vel(urf, uc, NULL, g);
  double R[4];
  for (int i = 0; i < 4; i++) R[i] = 4. / 3. * Erf * uc[0] * uc[i] + 1. / 3. * Erf * g->GG[0][i];
  o_lower(R, R, g->gg);
  for (int i = 0; i < 4; i++) uu[ee + i] = gdetu * R[i];
  return 0;
}

/* p2u: p2u.c:40-49 */
int o_p2u(const oracle_params *P, const double *pp, double *uu, const o_geom *g)
{
  o_p2u_mhd(P, pp, uu, g);
  if (P->rad) o_p2u_rad(P, pp, uu, g);
  return 0;
}

/* ---------------------------------------------------------------------------
 * MHD inversion: 1-D Newton in W = rho h gamma^2, u2p.c:1965-2352, residuals
 * f_u2p_hot u2p.c:733-814 (U2P_EQS_NOBLE) and f_u2p_entropy u2p.c:914-974.
 * cons = {Qn, Qt2, D, QdotBsq, Bsq, Sc, Qdotnp}
 * ------------------------------------------------------------------------- */
static void f_u2p_hot(const oracle_params *P, double Wp, const double *c, double *f, double *df, double *err)
{
  double G = P->gamma, GM1 = P->gamma - 1.;
  double Qn = c[0], Qtsq = c[1], D = c[2], QdotBsq = c[3], Bsq = c[4];
  double W = Wp + D;
  double X = Bsq + W;
  double Wsq = W * W;
  double W3 = Wsq * W;
  double X2 = X * X;
  double X3 = X2 * X;
  double v2 = (Wsq * Qtsq + QdotBsq * (Bsq + 2. * W)) / (Wsq * X2);
  double gamma2 = 1. / (1. - v2);
  double gamma = sqrt(gamma2);
  double wmrho0 = Wp / gamma2 - D * v2 / (1. + gamma);
  double u = wmrho0 / G;
  double p = (G - 1) * u;
  *f = Qn + W - p + 0.5 * Bsq * (1. + v2) - QdotBsq / 2. / Wsq;
  *err = fabs(*f) / (fabs(Qn) + fabs(W) + fabs(p) + fabs(0.5 * Bsq * (1. + v2)) + fabs(QdotBsq / 2. / Wsq));
  double dvsq = (-2.0 / X3 * (Qtsq + QdotBsq * (3.0 * W * X + Bsq * Bsq) / W3));
  double dp1 = (G - 1.) * (1. - v2) / G;
  double idwmrho0dp = GM1 / G;
  double dwmrho0dvsq = (D * (gamma * 0.5 - 1.0) - Wp);
  double drho0dvsq = -D * gamma * 0.5;
  double idrho0dp = 0.0;
  double dp2 = drho0dvsq * idrho0dp + dwmrho0dvsq * idwmrho0dp;
  double dpdW = dp1 + dp2 * dvsq;
  *df = 1. - dpdW + QdotBsq / (Wsq * W) + 0.5 * Bsq * dvsq;
}

static void f_u2p_entropy(const oracle_params *P, double Wp, const double *c, double *f, double *df, double *err)
{
  double G = P->gamma, GM1 = P->gamma - 1.;
  double Qtsq = c[1], D = c[2], QdotBsq = c[3], Bsq = c[4], Sc = c[5];
  double W = Wp + D;
  double X = Bsq + W;
  double Wsq = W * W;
  double W3 = Wsq * W;
  double X2 = X * X;
  double X3 = X2 * X;
  double v2 = (Wsq * Qtsq + QdotBsq * (Bsq + 2. * W)) / (Wsq * X2);
  double gamma2 = 1. / (1. - v2);
  double gamma = sqrt(gamma2);
  double rho0 = D / gamma;
  double wmrho0 = Wp / gamma2 - D * v2 / (1. + gamma);
  /* compute_specificentropy_wmrho0_idealgas, u2p.c:832-861 */
  double pressure = (GM1 / G) * wmrho0;
  double indexn = 1.0 / GM1;
  double inside = O_POW(pressure, indexn) / O_POW(rho0, indexn + 1.0);
  double Ss = O_LOG(inside);
  *f = -Sc + D * Ss;
  *err = fabs(*f) / (fabs(Sc) + fabs(D * Ss));
  double dSsdrho = G / ((1.0 - G) * rho0);
  double dSsdchi = 1.0 / (GM1 * wmrho0);
  double dwmrho0dW = 1.0 / gamma;
  double drho0dW = 0.0;
  double dwmrho0dvsq = (D * (gamma * 0.5 - 1.0) - Wp);
  double drho0dvsq = -D * gamma * 0.5;
  double dvsq = (-2.0 / X3 * (Qtsq + QdotBsq * (3.0 * W * X + Bsq * Bsq) / W3));
  double dSsdW = drho0dW * dSsdrho + dwmrho0dW * dSsdchi;
  double dSsdvsq = drho0dvsq * dSsdrho + dwmrho0dvsq * dSsdchi;
  double dSsdWp = dSsdW + dSsdvsq * dvsq;
  *df = D * dSsdWp;
}

static int w_out_of_bounds(double W, double Bsq, double QdotBsq, double Qtsq)
{
  return (W * W * W * (W + 2. * Bsq) - QdotBsq * (2. * W + Bsq)) <= W * W * (Qtsq - Bsq * Bsq);
}

/* u2p_solver_W from a given W start (Wstart > 0; the learned W guess, no
   reference counterpart) reporting its Newton iterations in *iters */
int o_u2p_solver_W_from(const oracle_params *P, const double *uu, double *pp, const o_geom *g, int etype,
                        double Wstart, int *iters);

int o_u2p_solver_W(const oracle_params *P, const double *uu, double *pp, const o_geom *g, int etype)
{
  return o_u2p_solver_W_from(P, uu, pp, g, etype, 0., NULL);
}

int o_u2p_solver_W_from(const oracle_params *P, const double *uu, double *pp, const o_geom *g, int etype,
                        double Wstart, int *iters)
{
  void (*fu)(const oracle_params *, double, const double *, double *, double *, double *) =
      (etype == O_U2P_ENTROPY) ? f_u2p_entropy : f_u2p_hot;
  double gdetu = g->gdet;
  double alpha = sqrt(-1. / g->GG[0][0]);
  double D = uu[0] / gdetu * alpha;
  double Sc = uu[5] / gdetu * alpha;
  double Qcov[4], Qcovp[4], Qcon[4], Qconp[4];
  Qcov[0] = (uu[1] / gdetu - uu[0] / gdetu) * alpha;
  Qcov[1] = uu[2] / gdetu * alpha;
  Qcov[2] = uu[3] / gdetu * alpha;
  Qcov[3] = uu[4] / gdetu * alpha;
  Qcovp[0] = uu[1] / gdetu * alpha;
  Qcovp[1] = uu[2] / gdetu * alpha;
  Qcovp[2] = uu[3] / gdetu * alpha;
  Qcovp[3] = uu[4] / gdetu * alpha;
  o_raise(Qcovp, Qconp, g->GG);
  o_raise(Qcov, Qcon, g->GG);
  double Bcon[4] = {0., 0., 0., 0.}, Bcov[4], Bsq = 0., QdotB = 0., QdotBsq = 0.;
  if (P->mag) {
    Bcon[1] = uu[6] / gdetu * alpha;
    Bcon[2] = uu[7] / gdetu * alpha;
    Bcon[3] = uu[8] / gdetu * alpha;
    o_lower(Bcon, Bcov, g->gg);
    Bsq = dot4(Bcon, Bcov);
    QdotB = dot4(Qcov, Bcon);
    QdotBsq = QdotB * QdotB;
  }
  double ncov[4] = {-alpha, 0., 0., 0.}, ncon[4];
  o_raise(ncov, ncon, g->GG);
  double Qn = Qcon[0] * ncov[0];
  double Qtcon[4], Qtcov[4];
  for (int i = 0; i < 4; i++) {
    Qtcon[i] = 0.;
    for (int j = 0; j < 4; j++) Qtcon[i] += ((i == j ? 1 : 0) + ncon[i] * ncov[j]) * Qcon[j];
  }
  o_lower(Qtcon, Qtcov, g->gg);
  double Qtsq = dot4(Qtcon, Qtcov);
  double bsqoasq = g->gg[0][1] * g->GG[0][1] + g->gg[0][2] * g->GG[0][2] + g->gg[0][3] * g->GG[0][3];
  double alphasq = alpha * alpha;
  double Dfactor = (-g->gttpert + alphasq * bsqoasq) / (alphasq + alpha);
  double Qdotnp = Qconp[0] * ncov[0] + D * (Dfactor);

  /* initial guess from the old primitives, u2p.c:2094-2112 */
  double utcon[4] = {0., pp[2], pp[3], pp[4]};
  double gamma2 = 1. + qsq_of(utcon, g);
  double gamma;
  double W = Wstart > 0. ? Wstart : (pp[0] + P->gamma * pp[1]) * gamma2;

  double cons[7] = {Qn, Qtsq, D, QdotBsq, Bsq, Sc, Qdotnp};
  double f0, dfdW, err;
  int i_increase = 0;
  if (iters) *iters = 0;
  for (;;) {
    f0 = dfdW = 0.;
    fu(P, W - D, cons, &f0, &dfdW, &err);
    if ((w_out_of_bounds(W, Bsq, QdotBsq, Qtsq) || !isfinite(f0) || !isfinite(dfdW)) && i_increase < 50) {
      W *= 10.;
      i_increase++;
      continue;
    }
    break;
  }
  if (i_increase >= 50) return -150;

  int iter = 0;
  double Wprev;
  do {
    Wprev = W;
    iter++;
    if (iters) *iters = iter;
    fu(P, W - D, cons, &f0, &dfdW, &err);
    if (err < P->u2pconv) break;
    if (dfdW == 0.) {
      W *= 1.1;
      continue;
    }
    double Wnew = W - f0 / dfdW;
    int idump = 0;
    double dumpfac = 1.;
    for (;;) {
      double f0t = 0., dft = 0., errt;
      fu(P, Wnew - D, cons, &f0t, &dft, &errt);
      if ((w_out_of_bounds(Wnew, Bsq, QdotBsq, Qtsq) || !isfinite(f0t) || !isfinite(dft)) && idump < 100) {
        idump++;
        dumpfac /= 2.;
        Wnew = W - dumpfac * f0 / dfdW;
        continue;
      }
      break;
    }
    if (idump >= 100) return -101;
    W = Wnew;
    if (fabs(W) > BIG) return -103;
    if (fabs((W - Wprev) / Wprev) < P->u2pconv && err < 1.e-1) break;
  } while (iter < 50);
  if (iter >= 50) return -102;
  if (!isfinite(W)) return -103;

  double Wsq = W * W;
  double Xsq = (Bsq + W) * (Bsq + W);
  double v2 = (Wsq * Qtsq + QdotBsq * (Bsq + 2. * W)) / (Wsq * Xsq);
  gamma2 = 1. / (1. - v2);
  gamma = sqrt(gamma2);
  double rho = D / gamma;
  double uint = 1. / P->gamma * (W / gamma2 - rho);
  for (int i = 1; i < 4; i++) utcon[i] = gamma / (W + Bsq) * (Qtcon[i] + QdotB * Bcon[i] / W);
  if (!isfinite(utcon[1])) return -120;
  if (uint < 0. || gamma2 < 0. || isnan(W) || !isfinite(W)) return -104;
  pp[0] = rho;
  pp[1] = uint;
  pp[2] = utcon[1];
  pp[3] = utcon[2];
  pp[4] = utcon[3];
  if (rho < 0.) return -105;
  pp[5] = o_calc_Sfromu(P, rho, uint);
  if (P->mag) {
    pp[6] = uu[6] / gdetu;
    pp[7] = uu[7] / gdetu;
    pp[8] = uu[8] / gdetu;
  }
  return 0;
}

/* u2p: u2p.c:176-485 (type 0: no entropy-decrease check) */
int o_u2p(const oracle_params *P, const double *uu0, double *pp, const o_geom *g,
          int corrected[3], int fixups[2])
{
  int nv = P->nv;
  double uu[O_NVMAX], ppbak[O_NVMAX];
  for (int i = 0; i < nv; i++) {
    uu[i] = uu0[i];
    ppbak[i] = pp[i];
  }
  double gdetu = g->gdet;
  int hdcorr = 0, radcor = 0;
  corrected[0] = corrected[1] = 0;
  fixups[0] = fixups[1] = 0;
  int ret = 0, u2pret = -1;
  if (uu[0] / gdetu < 0.) {
    pp[0] = P->rhofloor;
    uu[0] = P->rhofloor * gdetu;
    ret = -2;
    u2pret = -1;
  }
  u2pret = o_u2p_solver_W(P, uu, pp, g, O_U2P_HOT);
  if (u2pret < 0) {
    ret = -1;
    u2pret = o_u2p_solver_W(P, uu, pp, g, O_U2P_ENTROPY);
    if (u2pret < 0) ret = -2;
  }
  if (u2pret < 0) {
    ret = -3;
    for (int i = 0; i < nv; i++) pp[i] = ppbak[i];
  }
  if (ret < 0) hdcorr = 1;
  fixups[0] = (ret < -1) ? 1 : 0;
  corrected[2] = 0;
  if (P->rad) o_u2p_rad(P, uu, pp, g, &radcor);
  fixups[1] = (radcor > 0) ? 1 : 0;
  if (hdcorr > 0) corrected[0] = 1;
  if (radcor > 0) corrected[1] = 1;
  return ret;
}

/* check_floors_mhd: u2p.c:492-695 (ZAMO mass injection for b^2/rho) */
int o_check_floors_mhd(const oracle_params *P, double *pp, const o_geom *g)
{
  int ret = 0;
  double uu[O_NVMAX];
  o_p2u_mhd(P, pp, uu, g);
  if (pp[0] < P->rhofloor) {
    pp[0] = P->rhofloor;
    ret = -1;
  }
  if (pp[1] < P->uurhoratiomin * pp[0]) {
    pp[1] = P->uurhoratiomin * pp[0];
    ret = -1;
  }
  if (pp[1] > P->uurhoratiomax * pp[0]) {
    pp[1] = P->uurhoratiomax * pp[0];
    ret = -1;
  }
  if (P->mag) {
    double uc[4] = {0., pp[2], pp[3], pp[4]}, ucon[4], ucov[4], bcon[4], bcov[4];
    o_relr_to_4vel(uc, ucon, ucov, g);
    o_bcon(pp, ucon, ucov, bcon, 6);
    o_lower(bcon, bcov, g->gg);
    double magpre = dot4(bcon, bcov) / 2.;
    /* calc_normalobs_4vel relele.c:561-568, then VEL4->VELR with u^t known */
    double alp = 1. / sqrt(-g->GG[0][0]);
    double ncov[4] = {-alp, 0., 0., 0.}, etacon[4], etarel[4];
    o_raise(ncov, etacon, g->GG);
    etarel[0] = etacon[0];
    for (int i = 1; i < 4; i++) etarel[i] = etacon[i] - etarel[0] * g->GG[0][i] / g->GG[0][0];
    if (magpre > P->b2rhoratiomax * pp[0]) {
      double f = magpre / (P->b2rhoratiomax * pp[0]);
      double dpp[O_NVMAX], duu[O_NVMAX];
      double drho = pp[0] * (f - 1.);
      for (int i = 0; i < 9; i++) dpp[i] = 0.0;
      dpp[0] = drho;
      dpp[1] = pp[1] * drho / pp[0];
      dpp[2] = etarel[1];
      dpp[3] = etarel[2];
      dpp[4] = etarel[3];
      dpp[5] = 0.;
      dpp[6] = dpp[7] = dpp[8] = 0.;
      o_p2u_mhd(P, dpp, duu, g);
      for (int i = 0; i < 9; i++) uu[i] += duu[i];
      int rt = o_u2p_solver_W(P, uu, pp, g, O_U2P_HOT);
      if (rt < 0) rt = o_u2p_solver_W(P, uu, pp, g, O_U2P_ENTROPY);
      ret = -1;
    }
  }
  /* gamma ceiling (VELR), u2p.c:660-687 */
  {
    double qsq = 0.;
    for (int i = 1; i < 4; i++)
      for (int j = 1; j < 4; j++) qsq += pp[1 + i] * pp[1 + j] * g->gg[i][j];
    double gamma2 = 1. + qsq;
    if (gamma2 > P->gammamaxhd * P->gammamaxhd) {
      double qsqmax = P->gammamaxhd * P->gammamaxhd - 1.;
      double A = sqrt(qsqmax / qsq);
      for (int j = 1; j < 4; j++) pp[1 + j] *= A;
      ret = -1;
    }
  }
  if (ret < 0) pp[5] = o_calc_Sfromu(P, pp[0], pp[1]);
  return ret;
}
