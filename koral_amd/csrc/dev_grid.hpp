// Device grid description and the per-face / per-cell pieces of the explicit
// operator: reconstruction (avg2point finite.c:90-289), physical fluxes
// (f_flux_prime physics.c:687-820, f_flux_prime_rad rad.c:3278-3330), wave
// speeds (physics.c:10-241, rad.c:2451-2560) and metric source terms
// (physics.c:324-465).
#pragma once
#include "dev_implicit.hpp"

namespace kd {

enum { BC_COPY = 0, BC_PERIODIC = 1, BC_REFLECT = 2 };

// Ghost-inclusive box SX x SY x SZ, x fastest.  A collapsed dimension (n=1)
// has no ghosts.  Metric tables are SoA over the (x1, x2) plane.
struct DGrid {
  int nx, ny, nz, ng, ngx, ngy, ngz, sx, sy, sz;
  long ncell;
  long nfx, nfy, nfz;            // face counts: (sx+1)sy sz, sx(sy+1)sz, sx sy(sz+1)
  long npl, nplx, nply;          // metric plane sizes: sy*sx, sy*(sx+1), (sy+1)*sx
  const double *xb0, *xb1, *xb2; // face coordinates, index i+ng
  const double *gc, *Gc, *kr, *gfx, *Gfx, *gfy, *Gfy;
  int bc[6];
};

// Index functions of the ghost-inclusive box.  A collapsed dimension (n = 1)
// ignores its coordinate.  Cells span [-ng_d, n_d + ng_d) per dimension,
// faces of direction d one more along d.
KD_HD bool in_cells(int i, int n, int ng) { return n == 1 || (i >= -ng && i < n + ng); }
KD_HD bool in_faces(int i, int n, int ng) { return i >= -ng && i <= n + ng; }

KD_HD long cidx(const DGrid &G, int ix, int iy, int iz)
{
  KD_CHECK(in_cells(ix, G.nx, G.ngx) && in_cells(iy, G.ny, G.ngy) && in_cells(iz, G.nz, G.ngz), "cidx", ix, iy, iz);
  long y = G.ny > 1 ? iy + G.ngy : 0;
  long z = G.nz > 1 ? iz + G.ngz : 0;
  return (z * G.sy + y) * G.sx + ix + G.ngx;
}
KD_INL long fxidx(const DGrid &G, int ix, int iy, int iz)
{
  KD_CHECK(in_faces(ix, G.nx, G.ngx) && in_cells(iy, G.ny, G.ngy) && in_cells(iz, G.nz, G.ngz), "fxidx", ix, iy, iz);
  long y = G.ny > 1 ? iy + G.ngy : 0;
  long z = G.nz > 1 ? iz + G.ngz : 0;
  return (z * G.sy + y) * (G.sx + 1) + ix + G.ngx;
}
KD_INL long fyidx(const DGrid &G, int ix, int iy, int iz)
{
  KD_CHECK(G.ny > 1 && in_cells(ix, G.nx, G.ngx) && in_faces(iy, G.ny, G.ngy) && in_cells(iz, G.nz, G.ngz), "fyidx",
           ix, iy, iz);
  long z = G.nz > 1 ? iz + G.ngz : 0;
  return (z * (G.sy + 1) + iy + G.ngy) * G.sx + ix + G.ngx;
}
KD_INL long fzidx(const DGrid &G, int ix, int iy, int iz)
{
  KD_CHECK(G.nz > 1 && in_cells(ix, G.nx, G.ngx) && in_cells(iy, G.ny, G.ngy) && in_faces(iz, G.nz, G.ngz), "fzidx",
           ix, iy, iz);
  long y = G.ny > 1 ? iy + G.ngy : 0;
  return ((long)(iz + G.ngz) * G.sy + y) * G.sx + ix + G.ngx;
}
KD_INL long fidx(const DGrid &G, int d, int ix, int iy, int iz)
{
  return d == 0 ? fxidx(G, ix, iy, iz) : (d == 1 ? fyidx(G, ix, iy, iz) : fzidx(G, ix, iy, iz));
}
KD_INL long midx(const DGrid &G, int ix, int iy)
{
  KD_CHECK(in_cells(ix, G.nx, G.ngx) && in_cells(iy, G.ny, G.ngy), "midx", ix, iy, 0);
  return (long)(G.ny > 1 ? iy + G.ngy : 0) * G.sx + ix + G.ngx;
}
KD_HD int dimn(const DGrid &G, int d) { return d == 0 ? G.nx : (d == 1 ? G.ny : G.nz); }

// width of cell ic along d; the face coordinates xb_d hold n_d + 2 NG + 1
// entries in every dimension, collapsed ones included
KD_INL double size_x(const DGrid &G, int ic, int d)
{
  KD_CHECK(d >= 0 && d < 3 && ic >= -G.ng && ic < dimn(G, d) + G.ng, "size_x", ic, d, 0);
  const double *xb = d == 0 ? G.xb0 : (d == 1 ? G.xb1 : G.xb2);
  return xb[ic + 1 + G.ng] - xb[ic + G.ng];
}

// SoA element x[k + c] for a wave-uniform k (a variable's offset k = iv *
// stride) and a per-lane index c, with cb = 8 c.  On the device it is a
// buffer access: the base x + k forms a buffer resource in scalar registers
// and the 32-bit byte offset cb is the only vector operand, shared by every
// variable of the cell (a global access would carry a 64-bit address per
// load in vector registers).  Needs 8 c < 2^32 (the grid checks that every
// variable has fewer than 2^29 cells and faces, which also bounds the
// metric tables).
typedef unsigned kd_u32x2 __attribute__((ext_vector_type(2)));
KD_INL unsigned soa_off(long c) { return (unsigned)c * 8u; }
#if defined(__HIP_DEVICE_COMPILE__)
KD_INL __amdgpu_buffer_rsrc_t soa_rsrc(const double *x, long k)
{
  // raw buffer, no stride, the whole 32-bit range, 32-bit data format
  return __builtin_amdgcn_make_buffer_rsrc((void *)(x + k), (short)0, (int)0xffffffff, 0x00020000);
}
KD_INL double soa_ld(const double *x, long k, unsigned cb)
{
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(soa_rsrc(x, k), cb, 0, 0));
}
KD_INL void soa_st(double *x, long k, unsigned cb, double v)
{
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(kd_u32x2, v), soa_rsrc(x, k), cb, 0, 0);
}
#else
KD_INL double soa_ld(const double *x, long k, unsigned cb) { return *(const double *)((const char *)(x + k) + cb); }
KD_INL void soa_st(double *x, long k, unsigned cb, double v) { *(double *)((char *)(x + k) + cb) = v; }
#endif

KD_INL void geom_cell(const DGrid &G, int ix, int iy, Geom &g)
{
  const long m = midx(G, ix, iy), n = G.npl;
  const unsigned mb = soa_off(m);
  load_geom(g, [&](int c) { return soa_ld(G.gc, c * n, mb); }, [&](int c) { return soa_ld(G.Gc, c * n, mb); });
}

KD_INL void geom_face(const DGrid &G, int d, int ix, int iy, Geom &g)
{
  if (d == 0) {
    KD_CHECK(in_faces(ix, G.nx, G.ngx) && in_cells(iy, G.ny, G.ngy), "geom_face x", ix, iy, 0);
    const long m = (long)(G.ny > 1 ? iy + G.ngy : 0) * (G.sx + 1) + ix + G.ngx, n = G.nplx;
    const unsigned mb = soa_off(m);
    load_geom(g, [&](int c) { return soa_ld(G.gfx, c * n, mb); }, [&](int c) { return soa_ld(G.Gfx, c * n, mb); });
  } else if (d == 1) {
    KD_CHECK(G.ny > 1 && in_cells(ix, G.nx, G.ngx) && in_faces(iy, G.ny, G.ngy), "geom_face y", ix, iy, 0);
    const long m = (long)(iy + G.ngy) * G.sx + ix + G.ngx, n = G.nply;
    const unsigned mb = soa_off(m);
    load_geom(g, [&](int c) { return soa_ld(G.gfy, c * n, mb); }, [&](int c) { return soa_ld(G.Gfy, c * n, mb); });
  } else {
    geom_cell(G, ix, iy, g);
  }
}

// ---- reconstruction ----
KD_INL double mn(double a, double b) { return a < b ? a : b; }

KD_INL double minmod3(double a, double b, double c)
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

KD_INL double DMM(double X, double Y) { return .5 * (sgn(X) + sgn(Y)) * mn(fabs(X), fabs(Y)); }
KD_INL double DM4(double W, double X, double Y, double Z)
{
  return 0.125 * (sgn(W) + sgn(X)) * fabs((sgn(W) + sgn(Y)) * (sgn(W) + sgn(Z))) *
         mn(mn(fabs(W), fabs(X)), mn(fabs(Y), fabs(Z)));
}

// MP5 value at the + face of the centre of (um2, um1, u0, up1, up2)
KD_INL double mp5(double um2, double um1, double u0, double up1, double up2)
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

// PPM (finite.c:142-225) for one variable: returns both face values of the
// centre cell.  dx[0..4] are the widths of cells -2..+2.
KD_INL void ppm(const double *v, const double *dx, double &ul, double &ur)
{
  const double um2 = v[0], um1 = v[1], up1 = v[3], up2 = v[4];
  double u0 = v[2];
  const double dxm2 = dx[0], dxm1 = dx[1], dx0 = dx[2], dxp1 = dx[3], dxp2 = dx[4];
  double Z1 = (dxm1 + dx0) / (2. * dx0 + dxp1);
  double Z2 = (dxp2 + dxp1) / (2. * dxp1 + dx0);
  double DX = dxm1 + dx0 + dxp1 + dxp2;
  double dri = dx0 / (dxm1 + dx0 + dxp1) *
               ((2. * dxm1 + dx0) / (dxp1 + dx0) * (up1 - u0) + (dx0 + 2. * dxp1) / (dxm1 + dx0) * (u0 - um1));
  double drip1 = dxp1 / (dx0 + dxp1 + dxp2) *
                 ((2. * dx0 + dxp1) / (dxp2 + dxp1) * (up2 - up1) + (dxp1 + 2. * dxp2) / (dx0 + dxp1) * (up1 - u0));
  double drim1 = dxm1 / (dxm2 + dxm1 + dx0) *
                 ((2. * dxm2 + dxm1) / (dx0 + dxm1) * (u0 - um1) + (dxm1 + 2. * dx0) / (dxm2 + dxm1) * (um1 - um2));
  if ((up1 - u0) * (u0 - um1) > 0.)
    dri = mn(fabs(dri), mn(2. * fabs(u0 - um1), 2. * fabs(u0 - up1))) * sgn(dri);
  else
    dri = 0.;
  if ((up2 - up1) * (up1 - u0) > 0.)
    drip1 = mn(fabs(drip1), mn(2. * fabs(up1 - u0), 2. * fabs(up1 - up2))) * sgn(drip1);
  else
    drip1 = 0.;
  if ((u0 - um1) * (um1 - um2) > 0.)
    drim1 = mn(fabs(drim1), mn(2. * fabs(um1 - um2), 2. * fabs(um1 - u0))) * sgn(drim1);
  else
    drim1 = 0.;
  ur = u0 + dx0 / (dx0 + dxp1) * (up1 - u0) +
       1. / DX * ((2. * dxp1 * dx0) / (dxp1 + dx0) * (Z1 - Z2) * (up1 - u0) - dx0 * Z1 * drip1 + dxp1 * Z2 * dri);
  Z1 = (dxm2 + dxm1) / (2. * dxm1 + dx0);
  Z2 = (dxp1 + dx0) / (2. * dx0 + dxm1);
  DX = dxm2 + dxm1 + dx0 + dxp1;
  ul = um1 + dxm1 / (dxm1 + dx0) * (u0 - um1) +
       1. / DX * ((2. * dx0 * dxm1) / (dx0 + dxm1) * (Z1 - Z2) * (u0 - um1) - dxm1 * Z1 * dri + dx0 * Z2 * drim1);
  if ((ur - u0) * (u0 - ul) <= 0.) {
    ul = u0;
    ur = u0;
  }
  if ((ur - ul) * (ul - (3. * u0 - 2. * ur)) < 0.) ul = 3. * u0 - 2. * ur;
  if ((ur - ul) * ((3. * u0 - 2. * ul) - ur) < 0.) ur = 3. * u0 - 2. * ul;
}

// ---- physical fluxes ----
template <int MAG, int RAD>
KD_INL void flux_prime(const koral_params &P, const double *pp, int d, const Geom &g, double *ff)
{
  constexpr int EE = Lay<MAG, RAD>::EE;
  double gdetu = g.gdet;
  double ucon[4], ucov[4], bcon[4] = {0., 0., 0., 0.}, bcov[4] = {0., 0., 0., 0.}, bsq = 0.;
  relr_to_4vel(pp[2], pp[3], pp[4], ucon, ucov, g);
  if (MAG) {
    bcon4(pp[6], pp[7], pp[8], ucon, ucov, bcon);
    lower(bcon, bcov, g);
    bsq = dot4(bcon, bcov);
  }
  double rho = pp[0], u = pp[1], S = pp[5];
  double pre = (P.gamma - 1.) * u;
  double w = rho + u + pre;
  double eta = w + bsq;
  double etap = u + pre + bsq;
  double ptot = pre + 0.5 * bsq;
  // T^d_j = T^{dk} g_kj with T^{ik} = eta u^i u^k + ptot g^{ik} - b^i b^k
  double Tm[4];
  {
    double T[4];
#pragma unroll
    for (int k = 0; k < 4; k++) T[k] = eta * ucon[d] * ucon[k] + ptot * g.G[d][k] - bcon[d] * bcon[k];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      Tm[j] = 0.;
#pragma unroll
      for (int k = 0; k < 4; k++) Tm[j] += T[k] * g.g[k][j];
    }
  }
  double up1 = utp1(pp[2], pp[3], pp[4], g);
  ff[0] = gdetu * rho * ucon[d];
  ff[1] = gdetu * (etap * ucon[d] * ucov[0] + rho * ucon[d] * up1);
  if (MAG) ff[1] += -gdetu * bcon[d] * bcov[0];
  ff[2] = gdetu * (Tm[1]);
  ff[3] = gdetu * (Tm[2]);
  ff[4] = gdetu * (Tm[3]);
  ff[5] = gdetu * S * ucon[d];
  if (MAG) {
    ff[6] = gdetu * (bcon[1] * ucon[d] - bcon[d] * ucon[1]);
    ff[7] = gdetu * (bcon[2] * ucon[d] - bcon[d] * ucon[2]);
    ff[8] = gdetu * (bcon[3] * ucon[d] - bcon[d] * ucon[3]);
  }
  if (RAD) {
    double E = pp[EE], uc[4];
    relr_to_4vel(pp[EE + 1], pp[EE + 2], pp[EE + 3], uc, nullptr, g);
    double R[4];
#pragma unroll
    for (int k = 0; k < 4; k++) R[k] = 4. / 3. * E * uc[d] * uc[k] + 1. / 3. * E * g.G[d][k];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      double Rm = 0.;
#pragma unroll
      for (int k = 0; k < 4; k++) Rm += R[k] * g.g[k][j];
      ff[EE + j] = gdetu * (Rm);
    }
  }
}

// ---- wave speeds ----
// calc_wavespeeds_lr_core (gas: keeps cst1/cst2 order and leaves a0/a1
// untouched on a negative discriminant) and the radiative variant (min/max,
// negative discriminant clamped to zero)
KD_INL void ws_core(const double *ucon, const Geom &g, double w2, int dim, double &a0, double &a1,
                                        bool rad)
{
  double Acon[4], Bcon[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    double t = 0.;
#pragma unroll
    for (int k = 0; k < 4; k++) t += ((k == dim + 1) ? 1. : 0.) * g.G[i][k];
    Acon[i] = t;
    double s = 0.;
#pragma unroll
    for (int k = 0; k < 4; k++) s += ((k == 0) ? 1. : 0.) * g.G[i][k];
    Bcon[i] = s;
  }
  double Acov[4] = {0., 0., 0., 0.}, Bcov[4] = {1., 0., 0., 0.};
  Acov[dim + 1] = 1.;
  double Asq = dot4(Acon, Acov);
  double Bsq = dot4(Bcon, Bcov);
  double Au = dot4(Acov, ucon);
  double Bu = dot4(Bcov, ucon);
  double AB = dot4(Acon, Bcov);
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
    a0 = mn(c1, c2);
    a1 = mx(c1, c2);
  } else {
    a0 = c1;
    a1 = c2;
  }
}

// Per-cell record: ws[k] for k = 0..5 gas l/r per dim, 6..11 radiation l/r,
// 12..14 gas max |a| per dim, 15..17 radiation max (save_wavespeeds
// finite.c:296-398).  Without radiation the radiation slots are not stored.
constexpr int NWS = 18;

template <int MAG, int RAD>
KD_INL void wavespeeds_cell(const koral_params &P, const DGrid &G, const double *pp, int ix,
                                                int iy, int iz, const Geom &g, double *a)
{
  constexpr int EE = Lay<MAG, RAD>::EE;
  double ucon[4], ucov[4];
  relr_to_4vel(pp[2], pp[3], pp[4], ucon, ucov, g);
  double rho = pp[0], uu = pp[1];
  double pre = (P.gamma - 1.) * uu;
  double cs2 = P.gamma * pre / (rho + uu + pre);
  if (cs2 < 0.) cs2 = 0.;
  double va2 = 0.;
  if (MAG) {
    double bcon[4], bcov[4];
    bcon4(pp[6], pp[7], pp[8], ucon, ucov, bcon);
    lower(bcon, bcov, g);
    double bsq = dot4(bcon, bcov);
    double EF = rho + P.gamma * uu;
    double EEloc = bsq + EF;
    va2 = bsq / EEloc;
    if (va2 < 0.) va2 = 0.;
  }
  double vtot2 = cs2 + va2 - cs2 * va2;
  double a0 = 0., a1 = 0.;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    ws_core(ucon, g, vtot2, d, a0, a1, false);
    a[2 * d] = a0;
    a[2 * d + 1] = a1;
  }
#pragma unroll
  for (int k = 6; k < 12; k++) a[k] = 1.;
  if (RAD) {
    const int ii[3] = {ix, iy, iz};
    double chi = kappa_abs(P, rho, uu) + kappa_es(P, rho);
    double urf[4];
    relr_to_4vel(pp[EE + 1], pp[EE + 2], pp[EE + 3], urf, nullptr, g);
#pragma unroll
    for (int d = 0; d < 3; d++) {
      double s = sqrt(g.g[d + 1][d + 1]);
      double dxp = mx(size_x(G, ii[d], d) * s, size_x(G, ii[d] + 1, d) * s);
      double tau = chi * dxp;
      double rv2 = 1. / 3.;
      if (tau > 0.) {
        double rv2tau = 4. / 3. / tau * 4. / 3. / tau;
        rv2 = mn(rv2, rv2tau);
      }
      ws_core(urf, g, rv2, d, a[6 + 2 * d], a[7 + 2 * d], true);
    }
  }
#pragma unroll
  for (int k = 0; k < 6; k++) {
    if (a[2 * (k % 3) + 6 * (k / 3)] > 0.) a[2 * (k % 3) + 6 * (k / 3)] = 0.;
    if (a[2 * (k % 3) + 1 + 6 * (k / 3)] < 0.) a[2 * (k % 3) + 1 + 6 * (k / 3)] = 0.;
  }
#pragma unroll
  for (int d = 0; d < 3; d++) {
    a[12 + d] = mx(fabs(a[2 * d]), fabs(a[2 * d + 1]));
    a[15 + d] = mx(fabs(a[6 + 2 * d]), fabs(a[7 + 2 * d]));
  }
}

// ---- metric source terms (GDETIN=1), physics.c:324-465 ----
template <int MAG, int RAD>
KD_INL void metric_source(const koral_params &P, const DGrid &G, const double *pp, int ix, int iy,
                                              const Geom &g, double *ss)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  constexpr int EE = Lay<MAG, RAD>::EE;
  const long m = midx(G, ix, iy), npl = G.npl;
  const unsigned mb = soa_off(m);
  double gdetu = g.gdet;
  double ucon[4], ucov[4], bcon[4] = {0., 0., 0., 0.}, bsq = 0.;
  relr_to_4vel(pp[2], pp[3], pp[4], ucon, ucov, g);
  if (MAG) {
    double bcov[4];
    bcon4(pp[6], pp[7], pp[8], ucon, ucov, bcon);
    lower(bcon, bcov, g);
    bsq = dot4(bcon, bcov);
  }
  double pre = (P.gamma - 1.) * pp[1];
  double eta = pp[0] + pp[1] + pre + bsq;
  double ptot = pre + 0.5 * bsq;
#pragma unroll
  for (int iv = 0; iv < NV; iv++) ss[iv] = 0.;
  double ur4[4] = {0., 0., 0., 0.}, E = 0.;
  if (RAD) {
    E = pp[EE];
    relr_to_4vel(pp[EE + 1], pp[EE + 2], pp[EE + 3], ur4, nullptr, g);
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    // row k of T^k_l and R^k_l
    double Trow[4], Tm[4], Rm[4];
#pragma unroll
    for (int j = 0; j < 4; j++) Trow[j] = eta * ucon[k] * ucon[j] + ptot * g.G[k][j] - bcon[k] * bcon[j];
#pragma unroll
    for (int l = 0; l < 4; l++) {
      Tm[l] = 0.;
#pragma unroll
      for (int j = 0; j < 4; j++) Tm[l] += Trow[j] * g.g[j][l];
    }
    if (RAD) {
      double Rrow[4];
#pragma unroll
      for (int j = 0; j < 4; j++) Rrow[j] = 4. / 3. * E * ur4[k] * ur4[j] + 1. / 3. * E * g.G[k][j];
#pragma unroll
      for (int l = 0; l < 4; l++) {
        Rm[l] = 0.;
#pragma unroll
        for (int j = 0; j < 4; j++) Rm[l] += Rrow[j] * g.g[j][l];
      }
    }
#pragma unroll
    for (int l = 0; l < 4; l++) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        double K = soa_ld(G.kr, (long)(l * 16 + j * 4 + k) * npl, mb);
        ss[1 + j] += gdetu * Tm[l] * K;
        if (RAD) ss[EE + j] += gdetu * Rm[l] * K;
      }
    }
  }
}

}  // namespace kd
