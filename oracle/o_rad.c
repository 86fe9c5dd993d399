/*
 * Oracle: M1 radiation (stress tensor, four-force, inversion, floors) and the
 * per-cell implicit gas-radiation Newton solve.  Test infrastructure only.
 */
#include <math.h>
#include <float.h>
#include <string.h>
#include "oracle.h"
#include "o_math.h"

#define SMALL 1.e-50
#define BIG (1. / SMALL)
#define GAMMASMALLLIMIT (1.0 - 1E-10)
#define PI_KORAL 3.141592654 /* ko.h:47 */

static inline double mx(double a, double b) { return a > b ? a : b; }
static inline double sgn(double x) { return x > 0. ? 1. : (x < 0. ? -1. : 1.); } /* misc.c:86-92 */

static inline int ee_of(const oracle_params *P) { return 6 + 3 * P->mag; }

/* calc_Rij_M1: rad.c:1944-1975 -- R^{mu nu} */
void o_calc_Rij(const oracle_params *P, const double *pp, const o_geom *g, double R[4][4])
{
  int ee = ee_of(P);
  double E = pp[ee];
  double ur[4] = {0., pp[ee + 1], pp[ee + 2], pp[ee + 3]}, uc[4];
  o_relr_to_4vel(ur, uc, NULL, g);
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) R[i][j] = 4. / 3. * E * uc[i] * uc[j] + 1. / 3. * E * g->GG[i][j];
}

/* calc_PEQ_Tfromurho: physics.c:984-989 */
static double T_from_urho(const oracle_params *P, double u, double rho)
{
  double p = u * (P->gamma - 1.);
  return p / (P->k_boltz * rho / P->mu_gas / P->m_proton);
}

/* calc_kappa with the problem opacity (PROBLEMS/RADTUBE/kappa.c,
   PROBLEMS/KATOTORUS/kappa.c): absorption coefficient per unit length */
double o_kappa_abs(const oracle_params *P, const double *pp)
{
  double rho = pp[0];
  if (P->kappa_model == 0) return P->kappa_a_coef * rho;
  double T = T_from_urho(P, pp[1], rho);
  double rhocgs = rho / P->ggg * P->ccc * P->ccc / P->masscm / P->masscm;
  double Tcgs = T;
  double kcgs = P->kappa_a_coef * rhocgs / Tcgs / Tcgs / Tcgs / sqrt(Tcgs);
  return kcgs / P->ggg * P->ccc * P->ccc / P->masscm * rho;
}

/* calc_kappaes with PROBLEMS/KATOTORUS/kappaes.c form */
double o_kappa_es(const oracle_params *P, const double *pp) { return P->kappa_es_coef * pp[0]; }

/* calc_Gi: rad.c:1787-1888 -- contravariant radiative four-force */
void o_calc_Gi(const oracle_params *P, const double *pp, const o_geom *g, double Gi[4], int labframe)
{
  double R[4][4];
  o_calc_Rij(P, pp, g, R);
  double ur[4] = {0., pp[2], pp[3], pp[4]}, ucon[4], ucov[4];
  o_relr_to_4vel(ur, ucon, ucov, g);
  double rho = pp[0], u = pp[1];
  double p = (P->gamma - 1.) * u;
  double Tgas = p * P->mu_gas * P->m_proton / P->k_boltz / rho;
  double B = P->sigma_rad * O_POW(Tgas, 4.) / PI_KORAL;
  double kappa = o_kappa_abs(P, pp);
  double kes = o_kappa_es(P, pp);
  double Ruu = 0.;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) Ruu += R[i][j] * ucov[i] * ucov[j];
  for (int i = 0; i < 4; i++) {
    double Ru = 0.;
    for (int j = 0; j < 4; j++) Ru += R[i][j] * ucov[j];
    Gi[i] = -(kappa + kes) * Ru - (kes * Ruu + kappa * 4. * PI_KORAL * B) * ucon[i];
  }
  if (labframe == 0) Gi[0] = -kappa * 4. * PI_KORAL * B + kappa * Ruu;
}

/* calc_ff_Rtt: rad.c:2833-2862 -- Rtt = -Ehat and the gas u^mu */
void o_calc_ff_Rtt(const oracle_params *P, const double *pp, const o_geom *g, double *Rtt, double *ucon)
{
  double ur[4] = {0., pp[2], pp[3], pp[4]}, uc[4], uv[4];
  o_relr_to_4vel(ur, uc, uv, g);
  double R[4][4], Rm[4][4];
  o_calc_Rij(P, pp, g, R);
  /* indices_2221: R^i_j = R^{ik} g_{kj} */
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      Rm[i][j] = 0.;
      for (int k = 0; k < 4; k++) Rm[i][j] += R[i][k] * g->gg[k][j];
    }
  double r = 0.;
  for (int i1 = 0; i1 < 4; i1++)
    for (int i2 = 0; i2 < 4; i2++) r += -Rm[i1][i2] * uc[i2] * uv[i1];
  *Rtt = r;
  if (ucon)
    for (int i = 0; i < 4; i++) ucon[i] = uc[i];
}

/* calc_LTE_TfromE: rad.c:2430-2433 */
double o_TfromE(const oracle_params *P, double E) { return sqrt(sqrt((E / 4. / P->sigma_rad))); }

/* ---------------------------------------------------------------------------
 * M1 inversion R^t_mu -> (E_rf, u~^i): u2prad.c:108-481.
 * gamma^2=(u^t)^2 solves Q (4x+g)^2 = R^2 (g-8x), Q = R^t_mu R^{t mu},
 * R = R^{tt}, g = g^{tt}; the two roots are chosen as in get_m1closure_gammarel2.
 * ------------------------------------------------------------------------- */
static void m1_invariants(const double *Av, const o_geom *g, double *S, double *Sig)
{
  double s = g->GG[0][1] * Av[1] + g->GG[0][2] * Av[2] + g->GG[0][3] * Av[3];
  double q = 0.;
  for (int i = 1; i < 4; i++)
    for (int j = 1; j < 4; j++) q += g->GG[i][j] * Av[i] * Av[j];
  *S = s;
  *Sig = q;
}

/* get_m1closure_gammarel2_cold: u2prad.c:9-102 -- fixed R^t_i, prescribed
   gamma_rel^2; returns new R^t_t, E_rf and u~^i */
static void m1_cold(const o_geom *g, double *Avcov, double *Avcon, double gammarel2, double *Erf_out,
                    double *urf)
{
  double alpha = g->alpha;
  double x = gammarel2 / (alpha * alpha);
  double gtt = g->GG[0][0];
  double S, Sig;
  m1_invariants(Avcov, g, &S, &Sig);
  double SS = S * S - gtt * Sig;
  double tt = x * (gtt + x);
  double q4 = gtt + 4. * x;
  double rt = sqrt(SS * tt * q4 * q4);
  Avcov[0] = 0.25 * (-4. * S * tt + rt) / (gtt * tt);
  double Erf = 0.75 * rt / (tt * q4);
  o_raise(Avcov, Avcon, g->GG);
  double grel = sqrt(gammarel2);
  for (int j = 1; j < 4; j++)
    urf[j] = (Erf > 0.0) ? alpha * (Avcon[j] + 1. / 3. * Erf * g->GG[0][j] * (4.0 * gammarel2 - 1.0)) /
                               (4. / 3. * Erf * grel)
                         : 0.0;
  *Erf_out = Erf;
}

int o_u2p_rad(const oracle_params *P, const double *uu, double *pp, const o_geom *g, int *corrected)
{
  int ee = ee_of(P);
  double gdetu = g->gdet;
  *corrected = 0;
  double Avcov[4] = {uu[ee] / gdetu, uu[ee + 1] / gdetu, uu[ee + 2] / gdetu, uu[ee + 3] / gdetu};
  double Avcon[4];
  o_raise(Avcov, Avcon, g->GG);
  double alpha = g->alpha;

  /* get_m1closure_gammarel2, u2prad.c:108-227 */
  double gtt = g->GG[0][0];
  double S, Sig;
  m1_invariants(Avcov, g, &S, &Sig);
  double At = Avcov[0];
  double Q = gtt * At * At + 2. * At * S + Sig;
  double R = gtt * At + S;
  double sq = sqrt(R * R + 3. * Q * gtt);
  double base = Q * gtt + R * R;
  double gamma2 = -(base + R * sq) / (4. * Q);
  if (gamma2 < GAMMASMALLLIMIT || !isfinite(gamma2)) gamma2 = (-base + R * sq) / (4. * Q);
  double gammarel2 = gamma2 * alpha * alpha;
  if (gammarel2 > GAMMASMALLLIMIT && gammarel2 < 1.0) gammarel2 = 1.0;

  /* get_m1closure_Erf, u2prad.c:231-256 */
  double Erf = 3. * Avcon[0] * alpha * alpha / (4. * gammarel2 - 1.0);

  /* get_m1closure_urfconrel, u2prad.c:259-411 */
  double gmax = P->gammamaxrad;
  double urf[4];
  int nonfailure = gammarel2 >= 1.0 && Erf > P->eradfloor &&
                   gammarel2 <= gmax * gmax / GAMMASMALLLIMIT / GAMMASMALLLIMIT;
  if (nonfailure) {
    double grel = sqrt(gammarel2);
    for (int i = 0; i < 4; i++)
      urf[i] = alpha * (Avcon[i] + 1. / 3. * Erf * g->GG[0][i] * (4.0 * gammarel2 - 1.0)) / (4. / 3. * Erf * grel);
  } else {
    double Avs[4], Avsc[4], urs[4], Es;
    double Avf[4], Avfc[4], urff[4], Ef;
    for (int j = 0; j < 4; j++) {
      Avs[j] = Avf[j] = Avcov[j];
      urs[j] = urff[j] = 0.;
    }
    double g2slow = O_POW(1.0 + 10.0 * DBL_EPSILON, 2.0);
    double g2fast = gmax * gmax;
    m1_cold(g, Avs, Avsc, g2slow, &Es, urs);
    m1_cold(g, Avf, Avfc, g2fast, &Ef, urff);
    if (fabs(Avs[0] - Avcov[0]) > fabs(Avf[0] - Avcov[0])) {
      for (int j = 0; j < 4; j++) urf[j] = urff[j];
      gammarel2 = g2fast;
      Erf = Ef;
    } else {
      for (int j = 0; j < 4; j++) urf[j] = urs[j];
      gammarel2 = g2slow;
      Erf = Es;
    }
    *corrected = 1;
    if (!isfinite(Erf) || !isfinite(gammarel2) || !isfinite(urf[1]) || !isfinite(urf[2]) || !isfinite(urf[3]))
      return -1;
  }
  pp[ee] = Erf;
  pp[ee + 1] = urf[1];
  pp[ee + 2] = urf[2];
  pp[ee + 3] = urf[3];
  return 0;
}

/* check_floors_rad: u2prad.c:687-800 */
int o_check_floors_rad(const oracle_params *P, double *pp, const o_geom *g)
{
  int ee = ee_of(P);
  int ret = 0;
  double Rtt;
  o_calc_ff_Rtt(P, pp, g, &Rtt, NULL);
  double Ehat = -Rtt;
  double Eratio = pp[ee] / Ehat;
  if (pp[ee] < P->eradfloor) {
    pp[ee] = P->eradfloor;
    ret = -1;
  }
  if (Ehat < P->eerhoratiomin * pp[0]) {
    pp[ee] = Eratio * P->eerhoratiomin * pp[0];
    ret = -1;
  }
  if (Ehat > P->eerhoratiomax * pp[0]) {
    pp[0] = 1. / P->eerhoratiomax * Ehat;
    ret = -1;
  }
  if (Ehat < P->eeuuratiomin * pp[1]) {
    pp[ee] = Eratio * P->eeuuratiomin * pp[1];
    ret = -1;
  }
  if (Ehat > P->eeuuratiomax * pp[1]) {
    pp[1] = 1. / P->eeuuratiomax * Ehat;
    ret = -1;
  }
  return ret;
}

/* ---------------------------------------------------------------------------
 * Implicit lab-frame source: residual f_implicit_lab_4dprim rad.c:110-351
 * ------------------------------------------------------------------------- */
int o_f_implicit(const oracle_params *P, const double *ppin, const double *uu0, double *pp0, double dt,
                 const o_geom *g, double *f, const int *params, double *err0)
{
  int nv = P->nv, ee = ee_of(P);
  int whichprim = params[0], whicheq = params[1], whichframe = params[2];
  double gdetu = g->gdet;
  double uu[O_NVMAX], pp[O_NVMAX], err[4] = {0., 0., 0., 0.};
  int corr = 0, ret = 0, u2pret = 0;
  for (int i = 0; i < nv; i++) pp[i] = ppin[i];
  double ur[4] = {0., pp[2], pp[3], pp[4]}, ucon[4], ucov[4];
  o_relr_to_4vel(ur, ucon, ucov, g);
  if (whichprim == O_MHD) pp[0] = (uu0[0]) / gdetu / ucon[0];
  pp[5] = o_calc_Sfromu(P, pp[0], pp[1]);
  o_p2u(P, pp, uu, g);
  uu[5] = uu0[5] - (uu[ee] - uu0[ee]);
  if (whichprim == O_RAD) {
    uu[0] = uu0[0];
    uu[1] = uu0[1] - (uu[ee] - uu0[ee]);
    uu[2] = uu0[2] - (uu[ee + 1] - uu0[ee + 1]);
    uu[3] = uu0[3] - (uu[ee + 2] - uu0[ee + 2]);
    uu[4] = uu0[4] - (uu[ee + 3] - uu0[ee + 3]);
    int rt = o_u2p_solver_W(P, uu, pp, g, O_U2P_HOT);
    u2pret = (rt < 0) ? -2 : 0;
  }
  if (whichprim == O_MHD) {
    uu[ee] = uu0[ee] - (uu[1] - uu0[1]);
    uu[ee + 1] = uu0[ee + 1] - (uu[2] - uu0[2]);
    uu[ee + 2] = uu0[ee + 2] - (uu[3] - uu0[3]);
    uu[ee + 3] = uu0[ee + 3] - (uu[4] - uu0[4]);
    u2pret = o_u2p_rad(P, uu, pp, g, &corr);
  }
  if (corr != 0) ret = 1;
  if (u2pret < -1) return -1;

  double Gi[4];
  o_calc_Gi(P, pp, g, Gi, 1);
  o_lower(Gi, Gi, g->gg);
  if (whichprim == O_MHD) {
    for (int k = 1; k < 4; k++) {
      f[k] = uu[1 + k] - uu0[1 + k] - dt * gdetu * Gi[k];
      err[k] = (fabs(f[k]) > SMALL)
                   ? fabs(f[k]) / (fabs(uu[1 + k]) + fabs(uu0[1 + k]) + fabs(dt * gdetu * Gi[k]))
                   : 0.;
    }
  }
  if (whichprim == O_RAD) {
    for (int k = 1; k < 4; k++) {
      f[k] = uu[ee + k] - uu0[ee + k] + dt * gdetu * Gi[k];
      err[k] = (fabs(f[k]) > SMALL)
                   ? fabs(f[k]) / (fabs(uu[ee + k]) + fabs(uu0[ee + k]) + fabs(dt * gdetu * Gi[k]))
                   : 0.;
    }
  }
  if (whichframe == O_LAB) {
    if (whichprim == O_RAD) {
      int iv = (whicheq == O_ENERGYEQ) ? ee : 5;
      f[0] = uu[iv] - uu0[iv] + dt * gdetu * Gi[0];
      err[0] = (fabs(f[0]) > SMALL) ? fabs(f[0]) / (fabs(uu[iv]) + fabs(uu0[iv]) + fabs(dt * gdetu * Gi[0])) : 0.;
    } else {
      int iv = (whicheq == O_ENERGYEQ) ? 1 : 5;
      f[0] = uu[iv] - uu0[iv] - dt * gdetu * Gi[0];
      err[0] = (fabs(f[0]) > SMALL) ? fabs(f[0]) / (fabs(uu[iv]) + fabs(uu0[iv]) + fabs(dt * gdetu * Gi[0])) : 0.;
    }
  } else {
    /* fluid-frame energy/entropy equation, rad.c:273-337 */
    o_calc_Gi(P, pp, g, Gi, 0);
    double uconf[4], Rtt0, Rtt;
    o_calc_ff_Rtt(P, pp0, g, &Rtt0, uconf);
    o_calc_ff_Rtt(P, pp, g, &Rtt, uconf);
    double Ehat = -Rtt, Ehat0 = -Rtt0;
    double dtau = dt / uconf[0];
    if (whichprim == O_RAD) {
      if (whicheq == O_ENERGYEQ) {
        f[0] = Ehat - Ehat0 + dtau * Gi[0];
        err[0] = fabs(f[0]) / (fabs(Ehat) + fabs(Ehat0) + fabs(dtau * Gi[0]));
      } else {
        pp[5] = o_calc_Sfromu(P, pp[0], pp[1]);
        f[0] = pp[5] - pp0[5] - dtau * Gi[0];
        err[0] = fabs(f[0]) / (fabs(pp[5]) + fabs(pp0[5]) + fabs(dtau * Gi[0]));
      }
    } else {
      if (whicheq == O_ENERGYEQ) {
        f[0] = pp[1] - pp0[1] - dtau * Gi[0];
        err[0] = fabs(f[0]) / (fabs(pp[1]) + fabs(pp0[1]) + fabs(dtau * Gi[0]));
      } else {
        pp0[5] = o_calc_Sfromu(P, pp0[0], pp0[1]);
        pp[5] = o_calc_Sfromu(P, pp[0], pp[1]);
        f[0] = pp[5] - pp0[5] - dtau * Gi[0];
        err[0] = fabs(f[0]) / (fabs(pp[5]) + fabs(pp0[5]) + fabs(dtau * Gi[0]));
      }
    }
  }
  if (!isfinite(f[0]) || !isfinite(f[1]) || !isfinite(f[2]) || !isfinite(f[3])) return -1;
  *err0 = mx(mx(err[0], err[1]), mx(err[2], err[3]));
  return ret;
}

/* inverse_44matrix: misc.c:502-605 -- cofactor (adjugate) inverse; fails
   when an entry of the result is NaN */
int oracle_inverse44(const double *a, double *ia)
{
  /* 2x2 minors of rows 0,1 and rows 2,3 (Laplace expansion by pairs) */
  double s0 = a[0] * a[5] - a[4] * a[1];
  double s1 = a[0] * a[6] - a[4] * a[2];
  double s2 = a[0] * a[7] - a[4] * a[3];
  double s3 = a[1] * a[6] - a[5] * a[2];
  double s4 = a[1] * a[7] - a[5] * a[3];
  double s5 = a[2] * a[7] - a[6] * a[3];
  double c5 = a[10] * a[15] - a[14] * a[11];
  double c4 = a[9] * a[15] - a[13] * a[11];
  double c3 = a[9] * a[14] - a[13] * a[10];
  double c2 = a[8] * a[15] - a[12] * a[11];
  double c1 = a[8] * a[14] - a[12] * a[10];
  double c0 = a[8] * a[13] - a[12] * a[9];
  double det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
  double id = 1 / det;
  if (isnan(id)) return -1;
  double b[16];
  b[0] = (a[5] * c5 - a[6] * c4 + a[7] * c3);
  b[1] = (-a[1] * c5 + a[2] * c4 - a[3] * c3);
  b[2] = (a[13] * s5 - a[14] * s4 + a[15] * s3);
  b[3] = (-a[9] * s5 + a[10] * s4 - a[11] * s3);
  b[4] = (-a[4] * c5 + a[6] * c2 - a[7] * c1);
  b[5] = (a[0] * c5 - a[2] * c2 + a[3] * c1);
  b[6] = (-a[12] * s5 + a[14] * s2 - a[15] * s1);
  b[7] = (a[8] * s5 - a[10] * s2 + a[11] * s1);
  b[8] = (a[4] * c4 - a[5] * c2 + a[7] * c0);
  b[9] = (-a[0] * c4 + a[1] * c2 - a[3] * c0);
  b[10] = (a[12] * s4 - a[13] * s2 + a[15] * s0);
  b[11] = (-a[8] * s4 + a[9] * s2 - a[11] * s0);
  b[12] = (-a[4] * c3 + a[5] * c1 - a[6] * c0);
  b[13] = (a[0] * c3 - a[1] * c1 + a[2] * c0);
  b[14] = (-a[12] * s3 + a[13] * s1 - a[14] * s0);
  b[15] = (a[8] * s3 - a[9] * s1 + a[10] * s0);
  for (int i = 0; i < 16; i++) {
    ia[i] = b[i] * id;
    if (isnan(ia[i])) return -1;
  }
  return 0;
}

/* solve_implicit_lab_4dprim: rad.c:376-1218 -- damped Newton on 4 primitives
   with a one-sided finite-difference Jacobian.  x0 (may be NULL) replaces the
   start of the four iterated primitives (the reference starts from the old
   state); maxiter > 0 lowers RADIMPMAXITER for this attempt.  Both serve the
   learned initial guess, which has no reference counterpart. */
int o_solve_implicit_4dprim_from(const oracle_params *P, const double *uu00, const double *pp00, const o_geom *g,
                                 double dt, const int *params_in, double *pp, double global_time, int *iters_out,
                                 const double *xg, const double *xr, int maxiter)
{
  int nv = P->nv, ee = ee_of(P);
  int params[4] = {params_in[0], params_in[1], params_in[2], params_in[3]};
  double pp0[O_NVMAX], ppp[O_NVMAX], uu[O_NVMAX], uu0[O_NVMAX];
  double f1[4], f2[4], xxx[4], J[16], iJ[16], err, errbase;
  double gdetu = g->gdet;
  for (int i = 0; i < nv; i++) {
    uu0[i] = uu00[i];
    pp0[i] = pp00[i];
    uu[i] = uu0[i];
    pp[i] = pp0[i];
  }
  int whichprim = params[0];
  double EPS = P->radimpeps;
  double CONV = (params[1] == O_ENTROPYEQ) ? P->radimpentrconv : P->radimpconv;
  int MAXITER = P->radimpmaxiter;
  if (maxiter > 0 && maxiter < MAXITER) MAXITER = maxiter;
  int sh = (whichprim == O_MHD) ? 1 : ee;
  if (xg)
    for (int i = 0; i < 4; i++) pp[1 + i] = xg[i];
  if (xr && whichprim == O_RAD)
    for (int i = 0; i < 4; i++) pp[ee + i] = xr[i];
  int iter = 0, failed = 0;
  *iters_out = 0;
  for (;;) {
    iter++;
    for (int i = 0; i < nv; i++) ppp[i] = pp[i];
    if (o_f_implicit(P, pp, uu0, pp0, dt, g, f1, params, &err) < 0) {
      *iters_out = iter;
      return -1;
    }
    errbase = err;
    if (err < CONV) break;
    for (int j = 0; j < 4; j++) {
      double sign = -1.;
      for (;;) {
        double del;
        if (j == 0) {
          del = sign * EPS * ppp[sh];
          pp[sh] = ppp[sh] + del;
        } else {
          double veleps = EPS * sgn(ppp[j + sh]) * mx(1.e-6 / sqrt(g->gg[j][j]), fabs(ppp[j + sh]));
          del = (ppp[j + sh] >= 0.) ? sign * veleps : -sign * veleps;
          pp[j + sh] = ppp[j + sh] + del;
        }
        int fret = o_f_implicit(P, pp, uu0, pp0, dt, g, f2, params, &err);
        if (fret < 0) {
          if (sign > 0.) {
            failed = 1;
            break;
          }
          pp[j + sh] = ppp[j + sh];
          sign *= -1.;
          continue;
        }
        for (int i = 0; i < 4; i++) J[i * 4 + j] = (f2[i] - f1[i]) / (pp[j + sh] - ppp[j + sh]);
        pp[j + sh] = ppp[j + sh];
        break;
      }
      if (failed) break;
    }
    if (failed) break;
    if (oracle_inverse44(J, iJ) < 0) {
      failed = 1;
      break;
    }
    double xiapp = 1.;
    for (;;) {
      for (int i = 0; i < 4; i++) xxx[i] = ppp[i + sh];
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) xxx[i] -= xiapp * iJ[i * 4 + j] * f1[j];
      for (int i = 0; i < 4; i++) pp[i + sh] = xxx[i];
      if (whichprim == O_MHD) {
        double ur[4] = {0., pp[2], pp[3], pp[4]}, uc[4];
        o_relr_to_4vel(ur, uc, NULL, g);
        pp[0] = (uu0[0]) / gdetu / uc[0];
      }
      pp[5] = o_calc_Sfromu(P, pp[0], pp[1]);
      o_p2u(P, pp, uu, g);
      uu[5] = uu0[5] - (uu[ee] - uu0[ee]);
      int u2pret = 0;
      if (whichprim == O_RAD) {
        uu[0] = uu0[0];
        uu[1] = uu0[1] - (uu[ee] - uu0[ee]);
        uu[2] = uu0[2] - (uu[ee + 1] - uu0[ee + 1]);
        uu[3] = uu0[3] - (uu[ee + 2] - uu0[ee + 2]);
        uu[4] = uu0[4] - (uu[ee + 3] - uu0[ee + 3]);
        int rt = o_u2p_solver_W(P, uu, pp, g, O_U2P_HOT);
        u2pret = (rt < 0) ? -2 : 0;
      } else {
        int corr = 0;
        uu[ee] = uu0[ee] - (uu[1] - uu0[1]);
        uu[ee + 1] = uu0[ee + 1] - (uu[2] - uu0[2]);
        uu[ee + 2] = uu0[ee + 2] - (uu[3] - uu0[3]);
        uu[ee + 3] = uu0[ee + 3] - (uu[4] - uu0[4]);
        u2pret = o_u2p_rad(P, uu, pp, g, &corr);
        if (corr > 0) u2pret = -1;
      }
      double edenmin = ppp[sh] / P->maxenchangedown;
      double edenmax = ppp[sh] * P->maxenchangeup;
      if (xxx[0] > edenmin && xxx[0] < edenmax && u2pret >= -1) break;
      if (xxx[0] <= 0.) {
        xiapp *= ppp[sh] / (ppp[sh] + fabs(xxx[0]));
        xiapp *= 1.e-1;
      } else {
        xiapp /= P->dampingfactor;
      }
      /* rad.c:973; a NaN factor (0/0 when the iterate and the step are both
         zero) ends the attempt where the reference would never leave */
      if (!(xiapp >= P->maxradimpdamping)) {
        failed = 1;
        break;
      }
    }
    if (P->convrel_enabled && failed == 0 && global_time < 100.) {
      /* relative-change criterion under BHDISK_PROBLEMTYPE, rad.c:984-1027 */
      double f3[4];
      f3[0] = fabs((pp[sh] - ppp[sh]) / ppp[sh]);
      for (int i = 1; i < 4; i++) {
        f3[i] = pp[i + sh] - ppp[i + sh];
        f3[i] = fabs(f3[i] / mx(EPS, fabs(ppp[i + sh])));
      }
      double CR = P->radimpconvrel, CRE = P->radimpconvrelerr;
      if (params[1] == O_ENTROPYEQ) {
        CR = P->radimpconvrelentr;
        CRE = P->radimpconvrelentrerr;
      }
      if (f3[0] < CR && f3[1] < CR && f3[2] < CR && f3[3] < CR && errbase < CRE) break;
    }
    if (iter > MAXITER || failed == 1) break;
  }
  *iters_out = iter;
  if (iter > MAXITER || failed == 1) return -1;

  /* consistent final state, rad.c:1063-1140 */
  if (whichprim == O_MHD) {
    double ur[4] = {0., pp[2], pp[3], pp[4]}, uc[4];
    o_relr_to_4vel(ur, uc, NULL, g);
    pp[0] = (uu0[0]) / gdetu / uc[0];
  }
  pp[5] = o_calc_Sfromu(P, pp[0], pp[1]);
  o_p2u(P, pp, uu, g);
  int u2pret = 0;
  if (whichprim == O_RAD) {
    uu[0] = uu0[0];
    uu[1] = uu0[1] - (uu[ee] - uu0[ee]);
    uu[2] = uu0[2] - (uu[ee + 1] - uu0[ee + 1]);
    uu[3] = uu0[3] - (uu[ee + 2] - uu0[ee + 2]);
    uu[4] = uu0[4] - (uu[ee + 3] - uu0[ee + 3]);
    uu[5] = uu0[5];
    int rt = o_u2p_solver_W(P, uu, pp, g, O_U2P_HOT);
    u2pret = (rt < 0) ? -2 : 0;
  } else {
    int corr = 0;
    uu[ee] = uu0[ee] - (uu[1] - uu0[1]);
    uu[ee + 1] = uu0[ee + 1] - (uu[2] - uu0[2]);
    uu[ee + 2] = uu0[ee + 2] - (uu[3] - uu0[3]);
    uu[ee + 3] = uu0[ee + 3] - (uu[4] - uu0[4]);
    u2pret = o_u2p_rad(P, uu, pp, g, &corr);
    if (corr > 0) u2pret = -2;
  }
  if (u2pret < -1) return -1;
  return 0;
}

int o_solve_implicit_4dprim(const oracle_params *P, const double *uu00, const double *pp00, const o_geom *g,
                            double dt, const int *params_in, double *pp, double global_time, int *iters_out)
{
  return o_solve_implicit_4dprim_from(P, uu00, pp00, g, dt, params_in, pp, global_time, iters_out, NULL, NULL, 0);
}

/* solve_implicit_lab: rad.c:1225-1570, with implicit_lab_rad_source_term
   rad.c:2758-2795 flag handling.  uu/pp are updated in place on success. */
int o_solve_implicit_lab(const oracle_params *P, double *uu, double *pp, const o_geom *g, double dt,
                         double global_time, int *flags, o_imp_info *info)
{
  return o_solve_implicit_lab_guess(P, uu, pp, g, dt, global_time, flags, info, NULL, NULL, 0);
}

/* solve_implicit_lab preceded, when gxg is given, by one attempt (energy
   equation, lab frame) with at most maxiter iterations: from the gas start
   gxg on the gas primitives, or, when a radiation start gxr (E, F~^1..3) is
   given and the old state has E^ < RADIMPLICITTHRESHOLD u (the choice of
   attempt 0, rad.c:1381), from gxr on the radiation primitives with the gas
   part from gxg.  With gxr given, a cell that attempt does not solve tries
   the other primitive set from its start the same way.  A cell a learned try solves
   reports attempt 6, otherwise the unchanged six-attempt chain runs from the
   old state and total_iters includes the learned tries.  A NaN gas start
   (a cell the gate network leaves out) skips the learned tries.  The learned
   initial guess has no reference counterpart. */
int o_solve_implicit_lab_guess(const oracle_params *P, double *uu, double *pp, const o_geom *g, double dt,
                               double global_time, int *flags, o_imp_info *info, const double *gxg,
                               const double *gxr, int maxiter)
{
  int nv = P->nv;
  double pp00[O_NVMAX], uu00[O_NVMAX], ppn[O_NVMAX];
  for (int i = 0; i < nv; i++) {
    pp00[i] = pp[i];
    uu00[i] = uu[i];
  }
  if (flags) {
    flags[O_RADSOURCETYPEFLAG] = 3;
    flags[O_RADFIXUPFLAG] = 0;
  }
  double Rtt0;
  o_calc_ff_Rtt(P, pp00, g, &Rtt0, NULL);
  double Ehat = -Rtt0;
  int startwith = (Ehat < P->radimplicitthreshold * pp00[1]) ? O_RAD : O_MHD;
  int other = (startwith == O_RAD) ? O_MHD : O_RAD;
  /* fallback chain, rad.c:1389-1464 */
  const int seq[6][3] = {{0, O_ENERGYEQ, O_LAB}, {0, O_ENERGYEQ, O_FF}, {1, O_ENERGYEQ, O_LAB},
                         {1, O_ENERGYEQ, O_FF},  {0, O_ENTROPYEQ, O_LAB}, {1, O_ENTROPYEQ, O_LAB}};
  int ret = -1, total = 0, it = 0, a = 0;
  int params[4] = {0, 0, 0, 0};
  if (gxg && !isnan(gxg[0])) { /* a NaN gas start: no learned try */
    params[0] = gxr ? startwith : O_MHD;
    params[1] = O_ENERGYEQ;
    params[2] = O_LAB;
    ret = o_solve_implicit_4dprim_from(P, uu00, pp00, g, dt, params, ppn, global_time, &it, gxg, gxr, maxiter);
    total += it;
    if (ret != 0 && gxr) {
      params[0] = (params[0] == O_RAD) ? O_MHD : O_RAD;
      ret = o_solve_implicit_4dprim_from(P, uu00, pp00, g, dt, params, ppn, global_time, &it, gxg, gxr, maxiter);
      total += it;
    }
    if (ret == 0) a = 6;
  }
  for (; ret != 0 && a < 6; a++) {
    params[0] = seq[a][0] ? other : startwith;
    params[1] = seq[a][1];
    params[2] = seq[a][2];
    ret = o_solve_implicit_4dprim(P, uu00, pp00, g, dt, params, ppn, global_time, &it);
    total += it;
    if (ret == 0) break;
  }
  if (info) {
    info->total_iters = total;
    info->attempt = (ret == 0) ? a : -1;
    info->iters = (ret == 0) ? it : 0;
    info->whichprim = (ret == 0) ? params[0] : 0;
    info->whicheq = (ret == 0) ? params[1] : 0;
    info->whichframe = (ret == 0) ? params[2] : 0;
  }
  if (ret != 0) {
    if (flags) {
      flags[O_RADFIXUPFLAG] = -1;
      flags[O_RADIMPFIXUPFLAG] = -1;
    }
    return -1;
  }
  o_p2u(P, ppn, uu, g);
  for (int i = 0; i < nv; i++) pp[i] = ppn[i];
  if (flags) flags[O_RADIMPFIXUPFLAG] = 0;
  return 0;
}
