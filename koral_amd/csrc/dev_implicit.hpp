// Device implicit gas-radiation four-force solve (one lane per cell):
// residual f_implicit_lab_4dprim (rad.c:110-351), damped Newton
// solve_implicit_lab_4dprim (rad.c:376-1218) and the six-attempt fallback
// chain of solve_implicit_lab (rad.c:1225-1570).
#pragma once
#include "dev_cell.hpp"

namespace kd {

// f_implicit_lab_4dprim: returns -1 on failure, 1 if the M1 inversion was
// capped, 0 otherwise.  pp0 may be modified (fluid-frame entropy equation).
template <int MAG, int RAD, int WP, int WE, int WF>
KD_FUN int f_implicit(const koral_params &P, const double *ppin, const double *uu0, double *pp0, double dt,
                      const Geom &g, double *f, double &err0)
{
  constexpr int whichprim = WP, whicheq = WE, whichframe = WF;
  constexpr int NV = Lay<MAG, RAD>::NV;
  constexpr int EE = Lay<MAG, RAD>::EE;
  double gdetu = g.gdet;
  double uu[NV], pp[NV], err[4] = {0., 0., 0., 0.};
  int corr = 0, ret = 0, u2pret = 0;
#pragma unroll
  for (int i = 0; i < NV; i++) pp[i] = ppin[i];
  if (whichprim == P_MHD) {
    double ucon[4];
    relr_to_4vel(pp[2], pp[3], pp[4], ucon, nullptr, g);
    pp[0] = (uu0[0]) / gdetu / ucon[0];
  }
  pp[5] = Sfromu(P, pp[0], pp[1]);
  p2u<MAG, RAD>(P, pp, uu, g);
  uu[5] = uu0[5] - (uu[EE] - uu0[EE]);
  if (whichprim == P_RAD) {
    uu[0] = uu0[0];
    uu[1] = uu0[1] - (uu[EE] - uu0[EE]);
    uu[2] = uu0[2] - (uu[EE + 1] - uu0[EE + 1]);
    uu[3] = uu0[3] - (uu[EE + 2] - uu0[EE + 2]);
    uu[4] = uu0[4] - (uu[EE + 3] - uu0[EE + 3]);
    int rt = u2p_solver_W<MAG, RAD>(P, uu, pp, g, 0);
    u2pret = (rt < 0) ? -2 : 0;
  } else {
    uu[EE] = uu0[EE] - (uu[1] - uu0[1]);
    uu[EE + 1] = uu0[EE + 1] - (uu[2] - uu0[2]);
    uu[EE + 2] = uu0[EE + 2] - (uu[3] - uu0[3]);
    uu[EE + 3] = uu0[EE + 3] - (uu[4] - uu0[4]);
    u2pret = u2p_rad<MAG, RAD>(P, uu, pp, g, corr);
  }
  if (corr != 0) ret = 1;
  if (u2pret < -1) return -1;

  double Gi[4];
  calc_Gi<MAG, RAD>(P, pp, g, Gi, 1);
  lower(Gi, Gi, g);
  const int base = (whichprim == P_MHD) ? 1 : EE;
#pragma unroll
  for (int k = 1; k < 4; k++) {
    if (whichprim == P_MHD)
      f[k] = uu[base + k] - uu0[base + k] - dt * gdetu * Gi[k];
    else
      f[k] = uu[base + k] - uu0[base + k] + dt * gdetu * Gi[k];
    err[k] = (fabs(f[k]) > SMALL)
                 ? fabs(f[k]) / (fabs(uu[base + k]) + fabs(uu0[base + k]) + fabs(dt * gdetu * Gi[k]))
                 : 0.;
  }
  if (whichframe == LAB) {
    int iv = (whicheq == ENERGYEQ) ? base : 5;
    if (whichprim == P_RAD)
      f[0] = uu[iv] - uu0[iv] + dt * gdetu * Gi[0];
    else
      f[0] = uu[iv] - uu0[iv] - dt * gdetu * Gi[0];
    err[0] = (fabs(f[0]) > SMALL) ? fabs(f[0]) / (fabs(uu[iv]) + fabs(uu0[iv]) + fabs(dt * gdetu * Gi[0])) : 0.;
  } else {
    calc_Gi<MAG, RAD>(P, pp, g, Gi, 0);
    double uconf[4];
    double Rtt0 = ff_Rtt<MAG, RAD>(pp0, g, uconf);
    double Rtt = ff_Rtt<MAG, RAD>(pp, g, uconf);
    double Ehat = -Rtt, Ehat0 = -Rtt0;
    double dtau = dt / uconf[0];
    if (whichprim == P_RAD) {
      if (whicheq == ENERGYEQ) {
        f[0] = Ehat - Ehat0 + dtau * Gi[0];
        err[0] = fabs(f[0]) / (fabs(Ehat) + fabs(Ehat0) + fabs(dtau * Gi[0]));
      } else {
        pp[5] = Sfromu(P, pp[0], pp[1]);
        f[0] = pp[5] - pp0[5] - dtau * Gi[0];
        err[0] = fabs(f[0]) / (fabs(pp[5]) + fabs(pp0[5]) + fabs(dtau * Gi[0]));
      }
    } else {
      if (whicheq == ENERGYEQ) {
        f[0] = pp[1] - pp0[1] - dtau * Gi[0];
        err[0] = fabs(f[0]) / (fabs(pp[1]) + fabs(pp0[1]) + fabs(dtau * Gi[0]));
      } else {
        pp0[5] = Sfromu(P, pp0[0], pp0[1]);
        pp[5] = Sfromu(P, pp[0], pp[1]);
        f[0] = pp[5] - pp0[5] - dtau * Gi[0];
        err[0] = fabs(f[0]) / (fabs(pp[5]) + fabs(pp0[5]) + fabs(dtau * Gi[0]));
      }
    }
  }
  if (!isfinite(f[0]) || !isfinite(f[1]) || !isfinite(f[2]) || !isfinite(f[3])) return -1;
  err0 = mx(mx(err[0], err[1]), mx(err[2], err[3]));
  return ret;
}

// inverse_44matrix (misc.c:502-605): adjugate inverse, fails on NaN entries
KD_INL int inverse44(const double *a, double *ia)
{
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
  int bad = 0;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    ia[i] = b[i] * id;
    bad |= isnan(ia[i]) ? 1 : 0;
  }
  return bad ? -1 : 0;
}

// Newton-update helper: the "other half" of the state follows from
// conservation of U_gas+U_rad and an inversion (rad.c:880-950).
template <int MAG, int RAD, int WP>
KD_INL int complete_state(const koral_params &P, double *pp, double *uu, const double *uu0, const Geom &g,
                          bool final_pass)
{
  constexpr int whichprim = WP;
  constexpr int EE = Lay<MAG, RAD>::EE;
  double gdetu = g.gdet;
  if (whichprim == P_MHD) {
    double uc[4];
    relr_to_4vel(pp[2], pp[3], pp[4], uc, nullptr, g);
    pp[0] = (uu0[0]) / gdetu / uc[0];
  }
  pp[5] = Sfromu(P, pp[0], pp[1]);
  p2u<MAG, RAD>(P, pp, uu, g);
  if (!final_pass) uu[5] = uu0[5] - (uu[EE] - uu0[EE]);
  if (whichprim == P_RAD) {
    uu[0] = uu0[0];
    uu[1] = uu0[1] - (uu[EE] - uu0[EE]);
    uu[2] = uu0[2] - (uu[EE + 1] - uu0[EE + 1]);
    uu[3] = uu0[3] - (uu[EE + 2] - uu0[EE + 2]);
    uu[4] = uu0[4] - (uu[EE + 3] - uu0[EE + 3]);
    if (final_pass) uu[5] = uu0[5];
    int rt = u2p_solver_W<MAG, RAD>(P, uu, pp, g, 0);
    return (rt < 0) ? -2 : 0;
  }
  int corr = 0;
  uu[EE] = uu0[EE] - (uu[1] - uu0[1]);
  uu[EE + 1] = uu0[EE + 1] - (uu[2] - uu0[2]);
  uu[EE + 2] = uu0[EE + 2] - (uu[3] - uu0[3]);
  uu[EE + 3] = uu0[EE + 3] - (uu[4] - uu0[4]);
  int r = u2p_rad<MAG, RAD>(P, uu, pp, g, corr);
  if (corr > 0) r = final_pass ? -2 : -1;
  return r;
}

// One iteration of the damped Newton loop of solve_implicit_lab_4dprim
// (rad.c:376-1218) on the iterate pp (the attempt starts from pp = pp00,
// iter = 0).  Returns IMP_CONT to iterate again, IMP_OK with the final state
// in pp, or IMP_FAIL; iter counts the iterations of the attempt as the
// reference does.
enum { IMP_FAIL = -1, IMP_CONT = 0, IMP_OK = 1 };

template <int MAG, int RAD, int WP, int WE, int WF>
KD_FUN int imp_newton_step(const koral_params &P, const double *uu00, const double *pp00, const Geom &g, double dt,
                           double *pp, double global_time, int &iter)
{
  constexpr int whichprim = WP, whicheq = WE;
  constexpr int NV = Lay<MAG, RAD>::NV;
  constexpr int EE = Lay<MAG, RAD>::EE;
  double pp0[NV], ppp[NV], uu[NV];
  double f1[4], f2[4], xxx[4], J[16], iJ[16], err, errbase;
#pragma unroll
  for (int i = 0; i < NV; i++) pp0[i] = pp00[i];
  const double EPS = P.radimpeps;
  const double CONV = (whicheq == ENTROPYEQ) ? P.radimpentrconv : P.radimpconv;
  const int MAXITER = P.radimpmaxiter;
  const int sh = (whichprim == P_MHD) ? 1 : EE;
  int failed = 0;
  iter++;
#pragma unroll
  for (int i = 0; i < NV; i++) ppp[i] = pp[i];
  if (f_implicit<MAG, RAD, WP, WE, WF>(P, pp, uu00, pp0, dt, g, f1, err) < 0) return IMP_FAIL;
  errbase = err;
  if (err >= CONV) {
    // Jacobian by one-sided differences, column j from the iterate with
    // variable j + sh moved to xp (the loop over j is not unrolled, so the
    // arrays are addressed through selects and stay in registers)
    for (int j = 0; j < 4; j++) {
      double sign = -1.;
      double xj = ppp[sh];
#pragma unroll
      for (int k = 1; k < 4; k++) xj = (j == k) ? ppp[k + sh] : xj;
      const double gjj = (j == 1) ? g.g[1][1] : ((j == 2) ? g.g[2][2] : g.g[3][3]);
      for (;;) {
        double del;
        if (j == 0) {
          del = sign * EPS * xj;
        } else {
          double veleps = EPS * sgn(xj) * mx(1.e-6 / sqrt(gjj), fabs(xj));
          del = (xj >= 0.) ? sign * veleps : -sign * veleps;
        }
        const double xp = xj + del;
        double ppt[NV];
#pragma unroll
        for (int k = 0; k < NV; k++) ppt[k] = (k == j + sh) ? xp : pp[k];
        int fret = f_implicit<MAG, RAD, WP, WE, WF>(P, ppt, uu00, pp0, dt, g, f2, err);
        if (fret < 0) {
          if (sign > 0.) {
            failed = 1;
            break;
          }
          sign *= -1.;
          continue;
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const double d = (f2[i] - f1[i]) / (xp - xj);
#pragma unroll
          for (int k = 0; k < 4; k++)
            if (k == j) J[i * 4 + k] = d;
        }
        break;
      }
      if (failed) return IMP_FAIL;
    }
    if (inverse44(J, iJ) < 0) return IMP_FAIL;
    double xiapp = 1.;
    for (;;) {
#pragma unroll
      for (int i = 0; i < 4; i++) xxx[i] = ppp[i + sh];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) xxx[i] -= xiapp * iJ[i * 4 + j] * f1[j];
#pragma unroll
      for (int i = 0; i < 4; i++) pp[i + sh] = xxx[i];
      int u2pret = complete_state<MAG, RAD, WP>(P, pp, uu, uu00, g, false);
      double edenmin = ppp[sh] / P.maxenchangedown;
      double edenmax = ppp[sh] * P.maxenchangeup;
      if (xxx[0] > edenmin && xxx[0] < edenmax && u2pret >= -1) break;
      if (xxx[0] <= 0.) {
        xiapp *= ppp[sh] / (ppp[sh] + fabs(xxx[0]));
        xiapp *= 1.e-1;
      } else {
        xiapp /= P.dampingfactor;
      }
      // rad.c:973; a NaN factor (0/0 when the iterate and the step are both
      // zero) ends the attempt where the reference would never leave
      if (!(xiapp >= P.maxradimpdamping)) {
        failed = 1;
        break;
      }
    }
    bool conv_rel = false;
    if (P.convrel_enabled && failed == 0 && global_time < 100.) {
      double f3[4];
      f3[0] = fabs((pp[sh] - ppp[sh]) / ppp[sh]);
#pragma unroll
      for (int i = 1; i < 4; i++) {
        f3[i] = pp[i + sh] - ppp[i + sh];
        f3[i] = fabs(f3[i] / mx(EPS, fabs(ppp[i + sh])));
      }
      double CR = P.radimpconvrel, CRE = P.radimpconvrelerr;
      if (whicheq == ENTROPYEQ) {
        CR = P.radimpconvrelentr;
        CRE = P.radimpconvrelentrerr;
      }
      conv_rel = f3[0] < CR && f3[1] < CR && f3[2] < CR && f3[3] < CR && errbase < CRE;
    }
    if (!conv_rel) return (iter > MAXITER || failed == 1) ? IMP_FAIL : IMP_CONT;
  }
  // converged: the reference still rejects a solution past MAXITER
  if (iter > MAXITER) return IMP_FAIL;
  int u2pret = complete_state<MAG, RAD, WP>(P, pp, uu, uu00, g, true);
  return (u2pret < -1) ? IMP_FAIL : IMP_OK;
}

// imp_newton_step for attempt parameters known at run time: the six
// (primitives, equation, frame) kinds the fallback chain uses
template <int MAG, int RAD>
KD_INL int imp_newton_step_rt(const koral_params &P, const double *uu00, const double *pp00, const Geom &g,
                              double dt, int wp, int we, int wf, double *pp, double global_time, int &iter)
{
#define KD_STEP(A, B, C) imp_newton_step<MAG, RAD, A, B, C>(P, uu00, pp00, g, dt, pp, global_time, iter)
  if (wp == P_MHD) {
    if (we == ENERGYEQ) return wf == LAB ? KD_STEP(P_MHD, ENERGYEQ, LAB) : KD_STEP(P_MHD, ENERGYEQ, FF);
    if (wf == LAB) return KD_STEP(P_MHD, ENTROPYEQ, LAB);
  } else {
    if (we == ENERGYEQ) return wf == LAB ? KD_STEP(P_RAD, ENERGYEQ, LAB) : KD_STEP(P_RAD, ENERGYEQ, FF);
    if (wf == LAB) return KD_STEP(P_RAD, ENTROPYEQ, LAB);
  }
#undef KD_STEP
  return IMP_FAIL;  // the chain never asks for the fluid-frame entropy equation
}

// solve_implicit_lab_4dprim, rad.c:376-1218: one attempt from pp00
template <int MAG, int RAD>
KD_FUN int solve_4dprim(const koral_params &P, const double *uu00, const double *pp00, const Geom &g, double dt,
                        int whichprim, int whicheq, int whichframe, double *pp, double global_time, int &iters)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
#pragma unroll
  for (int i = 0; i < NV; i++) pp[i] = pp00[i];
  int iter = 0, st;
  do {
    st = imp_newton_step_rt<MAG, RAD>(P, uu00, pp00, g, dt, whichprim, whicheq, whichframe, pp, global_time, iter);
  } while (st == IMP_CONT);
  iters = iter;
  return st == IMP_OK ? 0 : -1;
}

// attempt a of the fallback chain (rad.c:1389-1464): energy LAB, energy FF,
// swapped LAB, swapped FF, entropy LAB, entropy swapped
KD_INL void imp_attempt(int a, int startwith, int &wp, int &we, int &wf)
{
  const int other = (startwith == P_RAD) ? P_MHD : P_RAD;
  wp = (a == 2 || a == 3 || a == 5) ? other : startwith;
  we = (a >= 4) ? ENTROPYEQ : ENERGYEQ;
  wf = (a == 1 || a == 3) ? FF : LAB;
}

// the fallback chain starts from RAD primitives when the fluid-frame
// radiation energy is small against the gas energy (rad.c:1340, 1381)
template <int MAG, int RAD>
KD_INL int imp_startwith(const koral_params &P, const double *pp00, const Geom &g)
{
  double Ehat = -ff_Rtt<MAG, RAD>(pp00, g, nullptr);
  return (Ehat < P.radimplicitthreshold * pp00[1]) ? P_RAD : P_MHD;
}

struct ImpInfo {
  int iters, attempt, whichprim, whicheq, whichframe, total;
};

// solve_implicit_lab + implicit_lab_rad_source_term (rad.c:1225-1570,
// rad.c:2758-2795).  Returns 0 on success with uu/pp updated.
template <int MAG, int RAD>
KD_FUN int solve_implicit_lab(const koral_params &P, double *uu, double *pp, const Geom &g, double dt,
                              double global_time, ImpInfo &info)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  double pp00[NV], ppn[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) pp00[i] = pp[i];
  const int startwith = imp_startwith<MAG, RAD>(P, pp00, g);
  int ret = -1, total = 0, it = 0, a = 0;
  int wp = startwith, we = ENERGYEQ, wf = LAB;
  for (a = 0; a < 6; a++) {
    imp_attempt(a, startwith, wp, we, wf);
    ret = solve_4dprim<MAG, RAD>(P, uu, pp00, g, dt, wp, we, wf, ppn, global_time, it);
    total += it;
    if (ret == 0) break;
  }
  info.total = total;
  info.attempt = (ret == 0) ? a : -1;
  info.iters = (ret == 0) ? it : 0;
  info.whichprim = (ret == 0) ? wp : 0;
  info.whicheq = (ret == 0) ? we : 0;
  info.whichframe = (ret == 0) ? wf : 0;
  if (ret != 0) return -1;
  p2u<MAG, RAD>(P, ppn, uu, g);
#pragma unroll
  for (int i = 0; i < NV; i++) pp[i] = ppn[i];
  return 0;
}

}  // namespace kd
