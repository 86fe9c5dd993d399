// Device (gfx950) per-point physics of KORAL's hot path: geometry algebra,
// primitive->conserved, the MHD and M1 inversions, floors, the radiative
// four-force and the implicit gas-radiation Newton solve.
//
// Everything is FP64 and register-resident; one lane owns one point.  State
// vectors are small fixed-size arrays indexed with compile-time constants after
// unrolling (template parameters MAG/RAD fix NV), so they live in VGPRs.
// Arithmetic order follows the reference expressions literally and the code is
// built with -ffp-contract=off; the transcendentals are the portable ones of
// kd_math.h, so a CPU build of the same arithmetic returns the same bits.
#pragma once
#include "dev_qual.hpp"
#include <math.h>
#include <float.h>
#include "koral_hip.h"
#include "kd_math.h"

namespace kd {

// Appends v to list for the lanes with pred (order within a wave by lane,
// across waves by arrival): the wave ballots its lanes, one lane reserves
// the wave's entries with one atomic add.  Every active lane of the wave
// calls it.
// Returns the entry's position for the lanes with pred, else -1.
#ifdef KD_HOST_CHECK
inline long append_index(bool pred, unsigned v, unsigned *list, unsigned *len)
{
  if (!pred) return -1;
  const unsigned slot = __atomic_fetch_add(len, 1u, __ATOMIC_RELAXED);
  list[slot] = v;
  return (long)slot;
}
#else
__device__ __forceinline__ long append_index(bool pred, unsigned v, unsigned *list, unsigned *len)
{
  const unsigned long long m = __ballot(pred);
  if (!m) return -1;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)m) - 1;
  unsigned base = 0;
  if (lane == leader) base = atomicAdd(len, (unsigned)__popcll(m));
  base = __shfl(base, leader);
  if (!pred) return -1;
  const unsigned slot = base + __popcll(m & ((1ull << lane) - 1ull));
  list[slot] = v;
  return (long)slot;
}
#endif


constexpr double SMALL = 1.e-50;
constexpr double BIG = 1. / SMALL;
constexpr double GAMMASMALLLIMIT = (1.0 - 1E-10);
constexpr double PI_KORAL = 3.141592654;  // ko.h:47

enum { RHO = 0, UU = 1, VX = 2, VY = 3, VZ = 4, ENTR = 5 };
enum { P_RAD = 1, P_MHD = 2 };
enum { ENERGYEQ = 0, ENTROPYEQ = 1, LAB = 0, FF = 1 };
enum { F_ENTROPY = 0, F_RADIMPFIXUP = 1, F_HDFIXUP = 2, F_RADFIXUP = 3, F_ENTROPY2 = 4,
       F_ENTROPY3 = 5, F_RADSOURCETYPE = 6, NFLAGS = 7 };

template <int MAG, int RAD>
struct Lay {
  static constexpr int NV = 6 + 3 * MAG + 4 * RAD;
  static constexpr int B1 = 6;
  static constexpr int EE = 6 + 3 * MAG;
};

struct Geom {
  double g[4][5];
  double G[4][5];
  double gdet, alpha, gttpert;
};

KD_INL double mx(double a, double b) { return a > b ? a : b; }
KD_INL double sgn(double x) { return x > 0. ? 1. : (x < 0. ? -1. : 1.); }

// fill_geometry, metric.c:1643-1676.  src(c) returns component c of the 20
// covariant numbers, srcG(c) of the 20 contravariant ones.
template <class F1, class F2>
KD_INL void load_geom(Geom &g, F1 src, F2 srcG)
{
#pragma unroll
  for (int c = 0; c < 20; c++) {
    g.g[c / 5][c % 5] = src(c);
    g.G[c / 5][c % 5] = srcG(c);
  }
  g.alpha = sqrt(-1. / g.G[0][0]);
  g.gdet = g.g[3][4];
  g.gttpert = g.G[3][4];
}

// indices_21 (frames.c:2086-2106) and indices_12 (frames.c:2060-2080)
KD_INL void lower(const double *a, double *b, const Geom &g)
{
  double t[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    t[i] = 0.;
#pragma unroll
    for (int k = 0; k < 4; k++) t[i] += a[k] * g.g[i][k];
  }
#pragma unroll
  for (int i = 0; i < 4; i++) b[i] = t[i];
}

KD_INL void raise_(const double *a, double *b, const Geom &g)
{
  double t[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    t[i] = 0.;
#pragma unroll
    for (int k = 0; k < 4; k++) t[i] += a[k] * g.G[i][k];
  }
#pragma unroll
  for (int i = 0; i < 4; i++) b[i] = t[i];
}

KD_INL double dot4(const double *a, const double *b)
{
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}

KD_INL double qsq_of(double v1, double v2, double v3, const Geom &g)
{
  const double v[4] = {0., v1, v2, v3};
  double q = 0.;
#pragma unroll
  for (int i = 1; i < 4; i++)
#pragma unroll
    for (int j = 1; j < 4; j++) q += v[i] * v[j] * g.g[i][j];
  return q;
}

// VELR -> VEL4 (conv_vels_both, relele.c:105-130; core relele.c:322-338)
KD_INL void relr_to_4vel(double v1, double v2, double v3, double *ucon, double *ucov,
                                             const Geom &g)
{
  double qsq = qsq_of(v1, v2, v3, g);
  double gamma2 = 1. + qsq;
  double alpha2 = -1. / g.G[0][0];
  double alpgam = sqrt(alpha2 * gamma2);
  ucon[0] = sqrt(gamma2 / alpha2);
  ucon[1] = v1 - alpgam * g.G[0][1];
  ucon[2] = v2 - alpgam * g.G[0][2];
  ucon[3] = v3 - alpgam * g.G[0][3];
  if (ucov) lower(ucon, ucov, g);
}

// calc_utp1 (VELR), p2u.c:548-575
KD_INL double utp1(double v1, double v2, double v3, const Geom &g)
{
  double qsq = qsq_of(v1, v2, v3, g);
  double gamma2 = 1. + qsq;
  double alphasq = -1. / g.G[0][0];
  double alpgam = sqrt(alphasq * gamma2);
  double bo = g.g[0][1] * g.G[0][1] + g.g[0][2] * g.G[0][2] + g.g[0][3] * g.G[0][3];
  double ud0 = 0.0;
  ud0 += v1 * g.g[0][1];
  ud0 += v2 * g.g[0][2];
  ud0 += v3 * g.g[0][3];
  return ud0 + (g.gttpert - alphasq * (bo + qsq)) / (1.0 + alpgam);
}

// calc_bcon_4vel, magn.c:7-23
KD_INL void bcon4(double B1, double B2, double B3, const double *ucon, const double *ucov,
                                      double *bcon)
{
  bcon[0] = B1 * ucov[1] + B2 * ucov[2] + B3 * ucov[3];
  bcon[1] = (B1 + bcon[0] * ucon[1]) / ucon[0];
  bcon[2] = (B2 + bcon[0] * ucon[2]) / ucon[0];
  bcon[3] = (B3 + bcon[0] * ucon[3]) / ucon[0];
}

// calc_Sfromu, physics.c:907-917
KD_INL double Sfromu(const koral_params &P, double rho, double u)
{
  double gm1 = P.gamma - 1.;
  double n = 1.0 / gm1;
  return rho * kd_log(kd_pow(gm1 * u, n) / kd_pow(rho, n + 1.0));
}

// p2u_mhd, p2u.c:56-159
template <int MAG, int RAD>
KD_INL void p2u_mhd(const koral_params &P, const double *pp, double *uu, const Geom &g)
{
  double gdetu = g.gdet;
  double rho = pp[0], u = pp[1];
  double ucon[4], ucov[4], bcon[4] = {0., 0., 0., 0.}, bcov[4] = {0., 0., 0., 0.}, bsq = 0.;
  double S = pp[5];
  relr_to_4vel(pp[2], pp[3], pp[4], ucon, ucov, g);
  if (MAG) {
    bcon4(pp[6], pp[7], pp[8], ucon, ucov, bcon);
    lower(bcon, bcov, g);
    bsq = dot4(bcon, bcov);
  }
  double ut = ucon[0];
  double rhout = rho * ut;
  double Sut = S * ut;
  double pre = (P.gamma - 1.) * u;
  double w = rho + u + pre;
  double eta = w + bsq;
  double etap = u + pre + bsq;
  double ptot = pre + 0.5 * bsq;
  double up1 = utp1(pp[2], pp[3], pp[4], g);
  double Tttt = etap * ucon[0] * ucov[0] + rho * ucon[0] * up1 + ptot - bcon[0] * bcov[0];
  double Ttr = eta * ucon[0] * ucov[1] - bcon[0] * bcov[1];
  double Ttth = eta * ucon[0] * ucov[2] - bcon[0] * bcov[2];
  double Ttph = eta * ucon[0] * ucov[3] - bcon[0] * bcov[3];
  uu[0] = gdetu * rhout;
  uu[1] = gdetu * Tttt;
  uu[2] = gdetu * Ttr;
  uu[3] = gdetu * Ttth;
  uu[4] = gdetu * Ttph;
  uu[5] = gdetu * Sut;
  if (MAG) {
    uu[6] = gdetu * pp[6];
    uu[7] = gdetu * pp[7];
    uu[8] = gdetu * pp[8];
  }
}

// p2u_rad (M1), p2u.c:253-307
template <int MAG, int RAD>
KD_INL void p2u_rad(const double *pp, double *uu, const Geom &g)
{
  constexpr int EE = Lay<MAG, RAD>::EE;
  double gdetu = g.gdet;
  double E = pp[EE];
  double uc[4];
  relr_to_4vel(pp[EE + 1], pp[EE + 2], pp[EE + 3], uc, nullptr, g);
  double R[4];
#pragma unroll
  for (int i = 0; i < 4; i++) R[i] = 4. / 3. * E * uc[0] * uc[i] + 1. / 3. * E * g.G[0][i];
  lower(R, R, g);
#pragma unroll
  for (int i = 0; i < 4; i++) uu[EE + i] = gdetu * R[i];
}

template <int MAG, int RAD>
KD_INL void p2u(const koral_params &P, const double *pp, double *uu, const Geom &g)
{
  p2u_mhd<MAG, RAD>(P, pp, uu, g);
  if (RAD) p2u_rad<MAG, RAD>(pp, uu, g);
}

// ---------------------------------------------------------------------------
// MHD inversion, u2p_solver_W (u2p.c:1965-2352).  Residuals f_u2p_hot
// (u2p.c:733-814) and f_u2p_entropy (u2p.c:914-974).
// ---------------------------------------------------------------------------
struct U2PCons {
  double Qn, Qtsq, D, QdotBsq, Bsq, Sc;
};

KD_INL void f_u2p(const koral_params &P, int etype, double Wp, const U2PCons &c, double &f,
                                      double &df, double &err)
{
  const double G = P.gamma, GM1 = P.gamma - 1.;
  const double D = c.D, Qtsq = c.Qtsq, QdotBsq = c.QdotBsq, Bsq = c.Bsq;
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
  double dvsq = (-2.0 / X3 * (Qtsq + QdotBsq * (3.0 * W * X + Bsq * Bsq) / W3));
  double dwmrho0dvsq = (D * (gamma * 0.5 - 1.0) - Wp);
  double drho0dvsq = -D * gamma * 0.5;
  if (etype == 0) {
    double u = wmrho0 / G;
    double p = (G - 1) * u;
    f = c.Qn + W - p + 0.5 * Bsq * (1. + v2) - QdotBsq / 2. / Wsq;
    err = fabs(f) / (fabs(c.Qn) + fabs(W) + fabs(p) + fabs(0.5 * Bsq * (1. + v2)) + fabs(QdotBsq / 2. / Wsq));
    double dp1 = (G - 1.) * (1. - v2) / G;
    double idwmrho0dp = GM1 / G;
    double idrho0dp = 0.0;
    double dp2 = drho0dvsq * idrho0dp + dwmrho0dvsq * idwmrho0dp;
    double dpdW = dp1 + dp2 * dvsq;
    df = 1. - dpdW + QdotBsq / (Wsq * W) + 0.5 * Bsq * dvsq;
  } else {
    double rho0 = D / gamma;
    double pressure = (GM1 / G) * wmrho0;
    double indexn = 1.0 / GM1;
    double inside = kd_pow(pressure, indexn) / kd_pow(rho0, indexn + 1.0);
    double Ss = kd_log(inside);
    f = -c.Sc + D * Ss;
    err = fabs(f) / (fabs(c.Sc) + fabs(D * Ss));
    double dSsdrho = G / ((1.0 - G) * rho0);
    double dSsdchi = 1.0 / (GM1 * wmrho0);
    double dwmrho0dW = 1.0 / gamma;
    double drho0dW = 0.0;
    double dSsdW = drho0dW * dSsdrho + dwmrho0dW * dSsdchi;
    double dSsdvsq = drho0dvsq * dSsdrho + dwmrho0dvsq * dSsdchi;
    double dSsdWp = dSsdW + dSsdvsq * dvsq;
    df = D * dSsdWp;
  }
}

KD_INL bool w_bad(double W, double Bsq, double QdotBsq, double Qtsq)
{
  return (W * W * W * (W + 2. * Bsq) - QdotBsq * (2. * W + Bsq)) <= W * W * (Qtsq - Bsq * Bsq);
}

// Wstart > 0 replaces the W start computed from pp (the learned W guess,
// which has no reference counterpart); iters, when given, receives the
// Newton iterations made.  budget > 0 (sequential path only) bounds the
// residual evaluations: the solver gives up with U2P_OVER_BUDGET, pp
// untouched, when a point needs more (a caller that can hand the point to
// the wave-parallel solver redoes it there from the start, same result).
//
// WAVE: a group of WAVE_GROUP lanes (an aligned part of a wave) runs the
// inversion of one point (uniform inputs within the group).  The two trial
// loops of the iteration -- W *= 10 until the residual is admissible (trials
// 0..49), and the step halving until the trial W is admissible (trials
// 0..99) -- are searches for the first admissible trial; the lanes of the
// group evaluate one trial each, in rounds of WAVE_GROUP trials (trial k is
// formed with the same k sequential multiplications, or as 2^-k, which k
// halvings of 1 give exactly), and take the first admissible one by a
// ballot restricted to the group, which is where the sequential loop stops,
// so the result is the same bit for bit.  A point whose iteration halves its
// step many times then costs a few evaluations of the residual instead of
// up to a hundred in sequence; the residual at the accepted trial also
// serves as the next iteration's residual.  Groups of 16 lanes (four points
// per wave) measured slower than whole waves on a torus whose deferred faces
// need long runs of step halvings (5.9 against 4.6 ms per step), so a group
// is the whole wave.
#ifdef KD_HOST_CHECK
#define KD_WAVE_SOLVER 0
#else
#define KD_WAVE_SOLVER 1
#endif
constexpr int WAVE_GROUP = 64;  // lanes per point of the WAVE solver (divides 64)
static_assert(64 % WAVE_GROUP == 0, "WAVE_GROUP divides the wave");
constexpr int U2P_OVER_BUDGET = -200;
template <int MAG, int RAD, bool WAVE = false, int WG = WAVE_GROUP>
KD_FUN int u2p_solver_W(const koral_params &P, const double *uu, double *pp, const Geom &g, int etype,
                        double Wstart = 0., int *iters = nullptr, int budget = 0)
{
  static_assert(64 % WG == 0, "a lane group divides the wave");
  double gdetu = g.gdet;
  double alpha = sqrt(-1. / g.G[0][0]);
  double D = uu[0] / gdetu * alpha;
  double Sc = uu[5] / gdetu * alpha;
  double Qcov[4], Qcon[4], Qconp0;
  Qcov[0] = (uu[1] / gdetu - uu[0] / gdetu) * alpha;
  Qcov[1] = uu[2] / gdetu * alpha;
  Qcov[2] = uu[3] / gdetu * alpha;
  Qcov[3] = uu[4] / gdetu * alpha;
  {
    double Qcovp[4] = {uu[1] / gdetu * alpha, Qcov[1], Qcov[2], Qcov[3]};
    double t = 0.;
#pragma unroll
    for (int k = 0; k < 4; k++) t += Qcovp[k] * g.G[0][k];
    Qconp0 = t;
  }
  raise_(Qcov, Qcon, g);
  double Bcon[4] = {0., 0., 0., 0.}, Bsq = 0., QdotB = 0., QdotBsq = 0.;
  if (MAG) {
    double Bcov[4];
    Bcon[1] = uu[6] / gdetu * alpha;
    Bcon[2] = uu[7] / gdetu * alpha;
    Bcon[3] = uu[8] / gdetu * alpha;
    lower(Bcon, Bcov, g);
    Bsq = dot4(Bcon, Bcov);
    QdotB = dot4(Qcov, Bcon);
    QdotBsq = QdotB * QdotB;
  }
  double ncov[4] = {-alpha, 0., 0., 0.}, ncon[4];
  raise_(ncov, ncon, g);
  double Qn = Qcon[0] * ncov[0];
  double Qtcon[4], Qtcov[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    Qtcon[i] = 0.;
#pragma unroll
    for (int j = 0; j < 4; j++) Qtcon[i] += ((i == j ? 1 : 0) + ncon[i] * ncov[j]) * Qcon[j];
  }
  lower(Qtcon, Qtcov, g);
  double Qtsq = dot4(Qtcon, Qtcov);
  U2PCons c{Qn, Qtsq, D, QdotBsq, Bsq, Sc};

  double gamma2 = 1. + qsq_of(pp[2], pp[3], pp[4], g);
  double W = Wstart > 0. ? Wstart : (pp[0] + P.gamma * pp[1]) * gamma2;
  double f0, dfdW, err;
  int i_increase = 0;
  int nev = 0;  // residual evaluations of the sequential path (budget)
  if (iters) *iters = 0;
#if KD_WAVE_SOLVER
  if constexpr (WAVE) {
    const int lane = threadIdx.x & 63, sub = lane % WG, g0 = lane - sub;
    const unsigned long long gmask = (WG == 64) ? ~0ull : (((1ull << WG) - 1ull) << g0);
    bool found = false;
    for (int r = 0; r * WG < 50 && !found; r++) {
      const int k = r * WG + sub;
      double Wk = W;
      for (int q = 0; q < k; q++) Wk *= 10.;
      bool ok = false;
      if (k < 50) {
        double f = 0., df = 0., e;
        f_u2p(P, etype, Wk - D, c, f, df, e);
        ok = !(w_bad(Wk, Bsq, QdotBsq, Qtsq) || !isfinite(f) || !isfinite(df));
      }
      const unsigned long long m = __ballot(ok) & gmask;
      if (m) {
        W = __shfl(Wk, __ffsll(m) - 1);
        found = true;
      }
    }
    if (!found) return -150;
  } else
#endif
  {
    for (;;) {
      f0 = dfdW = 0.;
      if (budget > 0 && ++nev > budget) return U2P_OVER_BUDGET;
      f_u2p(P, etype, W - D, c, f0, dfdW, err);
      if ((w_bad(W, Bsq, QdotBsq, Qtsq) || !isfinite(f0) || !isfinite(dfdW)) && i_increase < 50) {
        W *= 10.;
        i_increase++;
        continue;
      }
      break;
    }
    if (i_increase >= 50) return -150;
  }
  int iter = 0;
  double Wprev;
  // WAVE: the residual at the accepted trial, evaluated by the winning lane,
  // is the next iteration's residual at W (same inputs, same values).  The
  // sequential path evaluates it again: carrying it there lengthened
  // register lifetimes in the register-bound implicit kernels (3 % slower).
  bool have = false;
  double fc = 0., dfc = 0., ec = 0.;
  do {
    Wprev = W;
    iter++;
    if (iters) *iters = iter;
    if (WAVE && have) {
      f0 = fc;
      dfdW = dfc;
      err = ec;
    } else {
      if (budget > 0 && ++nev > budget) return U2P_OVER_BUDGET;
      f_u2p(P, etype, W - D, c, f0, dfdW, err);
    }
    have = false;
    if (err < P.u2pconv) break;
    if (dfdW == 0.) {
      W *= 1.1;
      continue;
    }
    double Wnew = W - f0 / dfdW;
#if KD_WAVE_SOLVER
    if constexpr (WAVE) {
      // trials k = 0..99 in rounds of WG; trial k halves the step k times
      const int lane = threadIdx.x & 63, sub = lane % WG, g0 = lane - sub;
      const unsigned long long gmask = (WG == 64) ? ~0ull : (((1ull << WG) - 1ull) << g0);
      int found = -1;
      for (int round = 0; round * WG < 100 && found < 0; round++) {
        const int k = round * WG + sub;
        // k halvings of 1 are exact: 2^-k
        const double dumpfac = ldexp(1., -k);
        const double Wt = k == 0 ? Wnew : W - dumpfac * f0 / dfdW;
        bool ok = false;
        double f0t = 0., dft = 0., errt = 0.;
        if (k < 100) {
          f_u2p(P, etype, Wt - D, c, f0t, dft, errt);
          ok = !(w_bad(Wt, Bsq, QdotBsq, Qtsq) || !isfinite(f0t) || !isfinite(dft));
        }
        const unsigned long long m = __ballot(ok) & gmask;
        if (m) {
          const int win = __ffsll(m) - 1;
          found = round * WG + win - g0;
          Wnew = __shfl(Wt, win);
          fc = __shfl(f0t, win);
          dfc = __shfl(dft, win);
          ec = __shfl(errt, win);
          have = true;
        }
      }
      if (found < 0) return -101;
    } else
#endif
    {
      int idump = 0;
      double dumpfac = 1.;
      for (;;) {
        double f0t = 0., dft = 0., errt;
        if (budget > 0 && ++nev > budget) return U2P_OVER_BUDGET;
        f_u2p(P, etype, Wnew - D, c, f0t, dft, errt);
        if ((w_bad(Wnew, Bsq, QdotBsq, Qtsq) || !isfinite(f0t) || !isfinite(dft)) && idump < 100) {
          idump++;
          dumpfac /= 2.;
          Wnew = W - dumpfac * f0 / dfdW;
          continue;
        }
        break;
      }
      if (idump >= 100) return -101;
    }
    W = Wnew;
    if (fabs(W) > BIG) return -103;
    if (fabs((W - Wprev) / Wprev) < P.u2pconv && err < 1.e-1) break;
  } while (iter < 50);
  if (iter >= 50) return -102;
  if (!isfinite(W)) return -103;

  double Wsq = W * W;
  double Xsq = (Bsq + W) * (Bsq + W);
  double v2 = (Wsq * Qtsq + QdotBsq * (Bsq + 2. * W)) / (Wsq * Xsq);
  gamma2 = 1. / (1. - v2);
  double gamma = sqrt(gamma2);
  double rho = D / gamma;
  double uint = 1. / P.gamma * (W / gamma2 - rho);
  double ut1 = gamma / (W + Bsq) * (Qtcon[1] + QdotB * Bcon[1] / W);
  double ut2 = gamma / (W + Bsq) * (Qtcon[2] + QdotB * Bcon[2] / W);
  double ut3 = gamma / (W + Bsq) * (Qtcon[3] + QdotB * Bcon[3] / W);
  if (!isfinite(ut1)) return -120;
  if (uint < 0. || gamma2 < 0. || isnan(W) || !isfinite(W)) return -104;
  pp[0] = rho;
  pp[1] = uint;
  pp[2] = ut1;
  pp[3] = ut2;
  pp[4] = ut3;
  if (rho < 0.) return -105;
  pp[5] = Sfromu(P, rho, uint);
  if (MAG) {
    pp[6] = uu[6] / gdetu;
    pp[7] = uu[7] / gdetu;
    pp[8] = uu[8] / gdetu;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// M1 closure inversion, u2prad.c:9-481
// ---------------------------------------------------------------------------
KD_INL void m1_inv(const double *Av, const Geom &g, double &S, double &Sig)
{
  S = g.G[0][1] * Av[1] + g.G[0][2] * Av[2] + g.G[0][3] * Av[3];
  double q = 0.;
#pragma unroll
  for (int i = 1; i < 4; i++)
#pragma unroll
    for (int j = 1; j < 4; j++) q += g.G[i][j] * Av[i] * Av[j];
  Sig = q;
}

// get_m1closure_gammarel2_cold, u2prad.c:9-102
KD_INL void m1_cold(const Geom &g, const double *Avc_in, double gammarel2, double &Av0,
                                        double &Erf, double *urf)
{
  double alpha = g.alpha;
  double x = gammarel2 / (alpha * alpha);
  double gtt = g.G[0][0];
  double S, Sig;
  m1_inv(Avc_in, g, S, Sig);
  double SS = S * S - gtt * Sig;
  double tt = x * (gtt + x);
  double q4 = gtt + 4. * x;
  double rt = sqrt(SS * tt * q4 * q4);
  Av0 = 0.25 * (-4. * S * tt + rt) / (gtt * tt);
  Erf = 0.75 * rt / (tt * q4);
  double Avcov[4] = {Av0, Avc_in[1], Avc_in[2], Avc_in[3]}, Avcon[4];
  raise_(Avcov, Avcon, g);
  double grel = sqrt(gammarel2);
#pragma unroll
  for (int j = 1; j < 4; j++)
    urf[j] = (Erf > 0.0) ? alpha * (Avcon[j] + 1. / 3. * Erf * g.G[0][j] * (4.0 * gammarel2 - 1.0)) /
                               (4. / 3. * Erf * grel)
                         : 0.0;
}

template <int MAG, int RAD>
KD_FUN int u2p_rad(const koral_params &P, const double *uu, double *pp, const Geom &g, int &corrected)
{
  constexpr int EE = Lay<MAG, RAD>::EE;
  double gdetu = g.gdet;
  corrected = 0;
  double Avcov[4] = {uu[EE] / gdetu, uu[EE + 1] / gdetu, uu[EE + 2] / gdetu, uu[EE + 3] / gdetu};
  double Avcon[4];
  raise_(Avcov, Avcon, g);
  double alpha = g.alpha;
  double gtt = g.G[0][0];
  double S, Sig;
  m1_inv(Avcov, g, S, Sig);
  double At = Avcov[0];
  double Q = gtt * At * At + 2. * At * S + Sig;
  double R = gtt * At + S;
  double sq = sqrt(R * R + 3. * Q * gtt);
  double base = Q * gtt + R * R;
  double gamma2 = -(base + R * sq) / (4. * Q);
  if (gamma2 < GAMMASMALLLIMIT || !isfinite(gamma2)) gamma2 = (-base + R * sq) / (4. * Q);
  double gammarel2 = gamma2 * alpha * alpha;
  if (gammarel2 > GAMMASMALLLIMIT && gammarel2 < 1.0) gammarel2 = 1.0;
  double Erf = 3. * Avcon[0] * alpha * alpha / (4. * gammarel2 - 1.0);
  double gmax = P.gammamaxrad;
  double u1, u2, u3;
  bool nonfailure = gammarel2 >= 1.0 && Erf > P.eradfloor &&
                    gammarel2 <= gmax * gmax / GAMMASMALLLIMIT / GAMMASMALLLIMIT;
  if (nonfailure) {
    double grel = sqrt(gammarel2);
    u1 = alpha * (Avcon[1] + 1. / 3. * Erf * g.G[0][1] * (4.0 * gammarel2 - 1.0)) / (4. / 3. * Erf * grel);
    u2 = alpha * (Avcon[2] + 1. / 3. * Erf * g.G[0][2] * (4.0 * gammarel2 - 1.0)) / (4. / 3. * Erf * grel);
    u3 = alpha * (Avcon[3] + 1. / 3. * Erf * g.G[0][3] * (4.0 * gammarel2 - 1.0)) / (4. / 3. * Erf * grel);
  } else {
    double g2slow = kd_pow(1.0 + 10.0 * DBL_EPSILON, 2.0);
    double g2fast = gmax * gmax;
    double As, Es, us[4] = {0., 0., 0., 0.}, Af, Ef, uf[4] = {0., 0., 0., 0.};
    m1_cold(g, Avcov, g2slow, As, Es, us);
    m1_cold(g, Avcov, g2fast, Af, Ef, uf);
    if (fabs(As - Avcov[0]) > fabs(Af - Avcov[0])) {
      u1 = uf[1]; u2 = uf[2]; u3 = uf[3];
      gammarel2 = g2fast;
      Erf = Ef;
    } else {
      u1 = us[1]; u2 = us[2]; u3 = us[3];
      gammarel2 = g2slow;
      Erf = Es;
    }
    corrected = 1;
    if (!isfinite(Erf) || !isfinite(gammarel2) || !isfinite(u1) || !isfinite(u2) || !isfinite(u3)) return -1;
  }
  pp[EE] = Erf;
  pp[EE + 1] = u1;
  pp[EE + 2] = u2;
  pp[EE + 3] = u3;
  return 0;
}

// ---------------------------------------------------------------------------
// radiation field and four-force
// ---------------------------------------------------------------------------
// calc_Rij_M1, rad.c:1944-1975
template <int MAG, int RAD>
KD_INL void Rij_M1(const double *pp, const Geom &g, double R[4][4])
{
  constexpr int EE = Lay<MAG, RAD>::EE;
  double E = pp[EE];
  double uc[4];
  relr_to_4vel(pp[EE + 1], pp[EE + 2], pp[EE + 3], uc, nullptr, g);
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) R[i][j] = 4. / 3. * E * uc[i] * uc[j] + 1. / 3. * E * g.G[i][j];
}

// calc_PEQ_Tfromurho, physics.c:984-989
KD_INL double T_from_urho(const koral_params &P, double u, double rho)
{
  double p = u * (P.gamma - 1.);
  return p / (P.k_boltz * rho / P.mu_gas / P.m_proton);
}

// calc_kappa with PROBLEMS/RADTUBE/kappa.c or PROBLEMS/KATOTORUS/kappa.c
KD_INL double kappa_abs(const koral_params &P, double rho, double u)
{
  if (P.kappa_model == 0) return P.kappa_a_coef * rho;
  double T = T_from_urho(P, u, rho);
  double rhocgs = rho / P.ggg * P.ccc * P.ccc / P.masscm / P.masscm;
  double Tcgs = T;
  double kcgs = P.kappa_a_coef * rhocgs / Tcgs / Tcgs / Tcgs / sqrt(Tcgs);
  return kcgs / P.ggg * P.ccc * P.ccc / P.masscm * rho;
}

KD_INL double kappa_es(const koral_params &P, double rho) { return P.kappa_es_coef * rho; }

// calc_Gi, rad.c:1787-1888
template <int MAG, int RAD>
KD_INL void calc_Gi(const koral_params &P, const double *pp, const Geom &g, double *Gi,
                                        int labframe)
{
  double R[4][4];
  Rij_M1<MAG, RAD>(pp, g, R);
  double ucon[4], ucov[4];
  relr_to_4vel(pp[2], pp[3], pp[4], ucon, ucov, g);
  double rho = pp[0], u = pp[1];
  double p = (P.gamma - 1.) * u;
  double Tgas = p * P.mu_gas * P.m_proton / P.k_boltz / rho;
  double B = P.sigma_rad * kd_pow(Tgas, 4.) / PI_KORAL;
  double kappa = kappa_abs(P, rho, u);
  double kes = kappa_es(P, rho);
  double Ruu = 0.;
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) Ruu += R[i][j] * ucov[i] * ucov[j];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    double Ru = 0.;
#pragma unroll
    for (int j = 0; j < 4; j++) Ru += R[i][j] * ucov[j];
    Gi[i] = -(kappa + kes) * Ru - (kes * Ruu + kappa * 4. * PI_KORAL * B) * ucon[i];
  }
  if (labframe == 0) Gi[0] = -kappa * 4. * PI_KORAL * B + kappa * Ruu;
}

// calc_ff_Rtt, rad.c:2833-2862
template <int MAG, int RAD>
KD_INL double ff_Rtt(const double *pp, const Geom &g, double *ucon_out)
{
  double uc[4], uv[4];
  relr_to_4vel(pp[2], pp[3], pp[4], uc, uv, g);
  double R[4][4];
  Rij_M1<MAG, RAD>(pp, g, R);
  double r = 0.;
#pragma unroll
  for (int i1 = 0; i1 < 4; i1++) {
#pragma unroll
    for (int i2 = 0; i2 < 4; i2++) {
      double Rm = 0.;
#pragma unroll
      for (int k = 0; k < 4; k++) Rm += R[i1][k] * g.g[k][i2];
      r += -Rm * uc[i2] * uv[i1];
    }
  }
  if (ucon_out)
#pragma unroll
    for (int i = 0; i < 4; i++) ucon_out[i] = uc[i];
  return r;
}

// ---------------------------------------------------------------------------
// high-level inversion and floors
// ---------------------------------------------------------------------------
// u2p, u2p.c:176-485 (type 0)
// With DEFER a state whose hot inversion fails is left for a later pass:
// the function returns U2P_DEFERRED at once and the caller redoes the state
// with DEFER false, so the entropy inversion is not inlined into the common
// path.  A failed hot inversion leaves pp as it was unless it failed with
// -105 (negative density after the update), so such a redo may start at the
// entropy inversion (SKIP_HOT) and reach the same state; U2P_DEFERRED_105
// marks the other cases, where the redo runs the hot inversion again: -105,
// and a hot inversion stopped by the budget (DEFER only, see u2p_solver_W).
constexpr int U2P_DEFERRED = 1, U2P_DEFERRED_105 = 2;
template <int MAG, int RAD, bool DEFER = false, bool SKIP_HOT = false, bool WAVE = false, int WG = WAVE_GROUP>
KD_FUN int u2p(const koral_params &P, const double *uu0, double *pp, const Geom &g, int *corrected, int *fixups,
               int budget = 0)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  double uu[NV], ppbak[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) {
    uu[i] = uu0[i];
    ppbak[i] = pp[i];
  }
  double gdetu = g.gdet;
  int radcor = 0;
  corrected[0] = corrected[1] = 0;
  fixups[0] = fixups[1] = 0;
  int ret = 0, u2pret = -1;
  if (uu[0] / gdetu < 0.) {
    pp[0] = P.rhofloor;
    uu[0] = P.rhofloor * gdetu;
    ret = -2;
  }
  u2pret = SKIP_HOT ? -1 : u2p_solver_W<MAG, RAD, WAVE, WG>(P, uu, pp, g, 0, 0., nullptr, DEFER ? budget : 0);
  if (u2pret < 0) {
    if (DEFER) return (u2pret == -105 || u2pret == U2P_OVER_BUDGET) ? U2P_DEFERRED_105 : U2P_DEFERRED;
    ret = -1;
    u2pret = u2p_solver_W<MAG, RAD, WAVE, WG>(P, uu, pp, g, 1);
    if (u2pret < 0) ret = -2;
  }
  if (u2pret < 0) {
    ret = -3;
#pragma unroll
    for (int i = 0; i < NV; i++) pp[i] = ppbak[i];
  }
  int hdcorr = ret < 0 ? 1 : 0;
  fixups[0] = (ret < -1) ? 1 : 0;
  corrected[2] = 0;
  if (RAD) u2p_rad<MAG, RAD>(P, uu, pp, g, radcor);
  fixups[1] = (radcor > 0) ? 1 : 0;
  if (hdcorr > 0) corrected[0] = 1;
  if (radcor > 0) corrected[1] = 1;
  return ret;
}

// check_floors_mhd, u2p.c:492-695.  With DEFER a state that needs the b^2/rho
// floor (ZAMO mass injection and re-inversion, u2p.c:600-634) is left for a
// later pass: the function returns FLOORS_DEFERRED at once, with pp partly
// floored, and the caller redoes the state with DEFER false.  Keeping the
// two inlined Newton inversions of that rare branch out of a kernel's
// common path is what DEFER is for.
constexpr int FLOORS_DEFERRED = 1;
template <int MAG, int RAD, bool DEFER = false, bool WAVE = false, int WG = WAVE_GROUP>
KD_FUN int floors_mhd(const koral_params &P, double *pp, const Geom &g)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  int ret = 0;
  double uu[NV];
  if (MAG) p2u_mhd<MAG, RAD>(P, pp, uu, g);
  if (pp[0] < P.rhofloor) {
    pp[0] = P.rhofloor;
    ret = -1;
  }
  if (pp[1] < P.uurhoratiomin * pp[0]) {
    pp[1] = P.uurhoratiomin * pp[0];
    ret = -1;
  }
  if (pp[1] > P.uurhoratiomax * pp[0]) {
    pp[1] = P.uurhoratiomax * pp[0];
    ret = -1;
  }
  if (MAG) {
    double ucon[4], ucov[4], bcon[4], bcov[4];
    relr_to_4vel(pp[2], pp[3], pp[4], ucon, ucov, g);
    bcon4(pp[6], pp[7], pp[8], ucon, ucov, bcon);
    lower(bcon, bcov, g);
    double magpre = dot4(bcon, bcov) / 2.;
    if (magpre > P.b2rhoratiomax * pp[0]) {
      if (DEFER) return FLOORS_DEFERRED;
      double alp = 1. / sqrt(-g.G[0][0]);
      double ncov[4] = {-alp, 0., 0., 0.}, etacon[4];
      raise_(ncov, etacon, g);
      double er1 = etacon[1] - etacon[0] * g.G[0][1] / g.G[0][0];
      double er2 = etacon[2] - etacon[0] * g.G[0][2] / g.G[0][0];
      double er3 = etacon[3] - etacon[0] * g.G[0][3] / g.G[0][0];
      double f = magpre / (P.b2rhoratiomax * pp[0]);
      double dpp[NV], duu[NV];
      double drho = pp[0] * (f - 1.);
      dpp[0] = drho;
      dpp[1] = pp[1] * drho / pp[0];
      dpp[2] = er1;
      dpp[3] = er2;
      dpp[4] = er3;
      dpp[5] = 0.;
      dpp[6] = dpp[7] = dpp[8] = 0.;
      p2u_mhd<MAG, RAD>(P, dpp, duu, g);
#pragma unroll
      for (int i = 0; i < 9; i++) uu[i] += duu[i];
      int rt = u2p_solver_W<MAG, RAD, WAVE, WG>(P, uu, pp, g, 0);
      if (rt < 0) rt = u2p_solver_W<MAG, RAD, WAVE, WG>(P, uu, pp, g, 1);
      ret = -1;
    }
  }
  {
    double qsq = 0.;
    const double v[4] = {0., pp[2], pp[3], pp[4]};
#pragma unroll
    for (int i = 1; i < 4; i++)
#pragma unroll
      for (int j = 1; j < 4; j++) qsq += v[i] * v[j] * g.g[i][j];
    double gamma2 = 1. + qsq;
    if (gamma2 > P.gammamaxhd * P.gammamaxhd) {
      double qsqmax = P.gammamaxhd * P.gammamaxhd - 1.;
      double A = sqrt(qsqmax / qsq);
      pp[2] *= A;
      pp[3] *= A;
      pp[4] *= A;
      ret = -1;
    }
  }
  if (ret < 0) pp[5] = Sfromu(P, pp[0], pp[1]);
  return ret;
}

// check_floors_rad, u2prad.c:687-800
template <int MAG, int RAD>
KD_FUN int floors_rad(const koral_params &P, double *pp, const Geom &g)
{
  constexpr int EE = Lay<MAG, RAD>::EE;
  int ret = 0;
  double Ehat = -ff_Rtt<MAG, RAD>(pp, g, nullptr);
  double Eratio = pp[EE] / Ehat;
  if (pp[EE] < P.eradfloor) {
    pp[EE] = P.eradfloor;
    ret = -1;
  }
  if (Ehat < P.eerhoratiomin * pp[0]) {
    pp[EE] = Eratio * P.eerhoratiomin * pp[0];
    ret = -1;
  }
  if (Ehat > P.eerhoratiomax * pp[0]) {
    pp[0] = 1. / P.eerhoratiomax * Ehat;
    ret = -1;
  }
  if (Ehat < P.eeuuratiomin * pp[1]) {
    pp[EE] = Eratio * P.eeuuratiomin * pp[1];
    ret = -1;
  }
  if (Ehat > P.eeuuratiomax * pp[1]) {
    pp[1] = 1. / P.eeuuratiomax * Ehat;
    ret = -1;
  }
  return ret;
}

// calc_primitives core (u2p.c:13-133) for an active, non-axis cell: returns
// the u2p code; sets flags and updates uu to follow the floored primitives.
// With DEFER a cell that needs the entropy inversion or the b^2/rho floor
// returns CP_DEFERRED_HOT (its hot inversion failed and left pp as it was)
// or CP_DEFERRED before writing uu (pp may be changed) and is redone by the
// caller with DEFER false, with SKIP_HOT for CP_DEFERRED_HOT.
constexpr int CP_DEFERRED = 100, CP_DEFERRED_HOT = 101;
template <int MAG, int RAD, bool DEFER = false, bool SKIP_HOT = false, bool WAVE = false, int WG = WAVE_GROUP>
KD_FUN int calc_primitives(const koral_params &P, double *uu, double *pp, const Geom &g, int &fl_entr,
                               int &fl_entr2, int &fl_hd, int &fl_rad, int budget = 0)
{
  int corrected[3] = {0, 0, 0}, fixups[2] = {0, 0};
  fl_entr = 0;
  fl_entr2 = 0;
  int r = u2p<MAG, RAD, DEFER, SKIP_HOT, WAVE, WG>(P, uu, pp, g, corrected, fixups, budget);
  if (DEFER && r == U2P_DEFERRED) return CP_DEFERRED_HOT;
  if (DEFER && r == U2P_DEFERRED_105) return CP_DEFERRED;
  if (corrected[0] == 1) fl_entr = 1;
  if (corrected[2] == 1) fl_entr2 = 1;
  if (floors_mhd<MAG, RAD, DEFER, WAVE, WG>(P, pp, g) == FLOORS_DEFERRED && DEFER) return CP_DEFERRED;
  if (RAD) floors_rad<MAG, RAD>(P, pp, g);
  p2u<MAG, RAD>(P, pp, uu, g);
  fl_hd = (fixups[0] > 0) ? 1 : 0;
  fl_rad = (fixups[1] > 0) ? -1 : 0;
  return r;
}

}  // namespace kd
