// Grid-level operators of one KORAL timestep on device-resident SoA state:
// ghost fill, wave speeds + time step, fused reconstruction / face floors /
// fluxes / HLL per face, flux-CT, conservative update with metric sources,
// u2p sweep with fixups, implicit sweep, and the RK2IMEX step body.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>
#include <vector>
#include "koral_hip.h"
#include "dev_grid.hpp"
#include "capi_util.hpp"

using namespace kd;

namespace {

constexpr int BS = 256;

struct Box {
  int x0, y0, z0, nx, ny, nz;
  __host__ __device__ long size() const { return (long)nx * ny * nz; }
  __device__ void at(long i, int &x, int &y, int &z) const
  {
    x = x0 + (int)(i % nx);
    long r = i / nx;
    y = y0 + (int)(r % ny);
    z = z0 + (int)(r / ny);
  }
};

inline dim3 blocks_for(long n) { return dim3((unsigned)((n + BS - 1) / BS)); }

template <int MAG, int RAD>
__device__ __forceinline__ void load_p(const double *p, long n, long c, double *v)
{
#pragma unroll
  for (int iv = 0; iv < Lay<MAG, RAD>::NV; iv++) v[iv] = p[iv * n + c];
}
template <int MAG, int RAD>
__device__ __forceinline__ void store_p(double *p, long n, long c, const double *v)
{
#pragma unroll
  for (int iv = 0; iv < Lay<MAG, RAD>::NV; iv++) p[iv * n + c] = v[iv];
}

// ---------------------------------------------------------------------------
// ghost fill (set_bc finite.c:2752-3700)
// ---------------------------------------------------------------------------
template <int MAG, int RAD>
__device__ __forceinline__ void p2u_at(const koral_params &P, const DGrid &G, const double *p, double *u, int ix,
                                       int iy, int iz)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  long c = cidx(G, ix, iy, iz);
  Geom g;
  geom_cell(G, ix, iy, g);
  double pp[NV], uu[NV];
  load_p<MAG, RAD>(p, G.ncell, c, pp);
  kd::p2u<MAG, RAD>(P, pp, uu, g);
  store_p<MAG, RAD>(u, G.ncell, c, uu);
}

template <int MAG, int RAD>
__global__ void __launch_bounds__(BS) k_p2u_box(koral_params P, DGrid G, const double *p, double *u, Box b)
{
  long i = (long)blockIdx.x * BS + threadIdx.x;
  if (i >= b.size()) return;
  int ix, iy, iz;
  b.at(i, ix, iy, iz);
  p2u_at<MAG, RAD>(P, G, p, u, ix, iy, iz);
}

template <int MAG, int RAD>
__global__ void __launch_bounds__(BS) k_bc_faces(koral_params P, DGrid G, double *p, double *u, Box b)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  long i = (long)blockIdx.x * BS + threadIdx.x;
  if (i >= b.size()) return;
  int ii[3];
  b.at(i, ii[0], ii[1], ii[2]);
  int d = -1, hi = 0, out = 0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    int n = dimn(G, k);
    if (n > 1 && ii[k] < 0) { d = k; hi = 0; out++; }
    if (n > 1 && ii[k] >= n) { d = k; hi = 1; out++; }
  }
  if (out != 1) return;
  int src[3] = {ii[0], ii[1], ii[2]};
  int n = dimn(G, d);
  int type = G.bc[2 * d + hi];
  bool flip = false;
  if (type == BC_PERIODIC) {
    if (src[d] < 0) src[d] += n;
    if (src[d] > n - 1) src[d] -= n;
  } else if (type == BC_REFLECT) {
    src[d] = hi ? 2 * n - src[d] - 1 : -src[d] - 1;
    flip = true;
  } else {
    src[d] = hi ? n - 1 : 0;
  }
  double pv[NV];
  load_p<MAG, RAD>(p, G.ncell, cidx(G, src[0], src[1], src[2]), pv);
  if (flip) {
    pv[2 + d] = -pv[2 + d];
    if (MAG) pv[6 + d] = -pv[6 + d];
    if (RAD) pv[Lay<MAG, RAD>::EE + 1 + d] = -pv[Lay<MAG, RAD>::EE + 1 + d];
  }
  long c = cidx(G, ii[0], ii[1], ii[2]);
  store_p<MAG, RAD>(p, G.ncell, c, pv);
  Geom g;
  geom_cell(G, ii[0], ii[1], g);
  double uu[NV];
  kd::p2u<MAG, RAD>(P, pv, uu, g);
  store_p<MAG, RAD>(u, G.ncell, c, uu);
}

// one thread per (corner of the a-b edge, position along the third dim):
// finite.c:3336-3600 recipe (copies of the far edge ghosts, two diagonal
// averages), executed in the reference's order
template <int MAG, int RAD>
__global__ void __launch_bounds__(BS) k_bc_edge(koral_params P, DGrid G, double *p, double *u, int a, int b)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  const int NG = G.ng;
  const int c = 3 - a - b;
  const int na = dimn(G, a), nb = dimn(G, b), nc = dimn(G, c);
  const int ngc = c == 0 ? G.ngx : (c == 1 ? G.ngy : G.ngz);  // 0 if collapsed
  const int nk = nc + 2 * ngc;
  long t = (long)blockIdx.x * BS + threadIdx.x;
  if (t >= 4L * nk) return;
  const int corner = (int)(t / nk);
  const int k = -ngc + (int)(t % nk);
  const int ahi = corner >> 1, bhi = corner & 1;
  const int ea = ahi ? na : -1, ia = ahi ? na - 1 : 0, ea2 = ahi ? na + 1 : -2;
  const int eb = bhi ? nb : -1, ib = bhi ? nb - 1 : 0, eb2 = bhi ? nb + 1 : -2;
  auto cell = [&](int va, int vb) {
    int q[3];
    q[a] = va;
    q[b] = vb;
    q[c] = k;
    return cidx(G, q[0], q[1], q[2]);
  };
  auto fill_u = [&](int va, int vb) {
    int q[3];
    q[a] = va;
    q[b] = vb;
    q[c] = k;
    p2u_at<MAG, RAD>(P, G, p, u, q[0], q[1], q[2]);
  };
  for (int i = 0; i < NG - 1; i++) {
    int ga = ahi ? na + i + 1 : -NG + i;
    int gb = bhi ? nb + i + 1 : -NG + i;
    long d1 = cell(ga, eb), s1 = cell(ga, ib), d2 = cell(ea, gb), s2 = cell(ia, gb);
#pragma unroll
    for (int iv = 0; iv < NV; iv++) p[iv * G.ncell + d1] = p[iv * G.ncell + s1];
#pragma unroll
    for (int iv = 0; iv < NV; iv++) p[iv * G.ncell + d2] = p[iv * G.ncell + s2];
    fill_u(ga, eb);
    fill_u(ea, gb);
  }
  {
    long dd = cell(ea, eb), aa = cell(ea, ib), bb = cell(ia, eb);
#pragma unroll
    for (int iv = 0; iv < NV; iv++) p[iv * G.ncell + dd] = .5 * (p[iv * G.ncell + aa] + p[iv * G.ncell + bb]);
    fill_u(ea, eb);
  }
  {
    long dd = cell(ea2, eb2), aa = cell(ea2, eb), bb = cell(ea, eb2);
#pragma unroll
    for (int iv = 0; iv < NV; iv++) p[iv * G.ncell + dd] = .5 * (p[iv * G.ncell + aa] + p[iv * G.ncell + bb]);
    fill_u(ea2, eb2);
  }
}

template <int MAG, int RAD>
__global__ void __launch_bounds__(BS) k_bc_total(koral_params P, DGrid G, double *p, double *u)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  const int NG = G.ng;
  long per = (long)NG * NG * NG;
  long t = (long)blockIdx.x * BS + threadIdx.x;
  if (t >= 8 * per) return;
  int blk = (int)(t / per);
  int r = (int)(t % per);
  int hx = blk & 1, hy = (blk >> 1) & 1, hz = (blk >> 2) & 1;
  int ix = (hx ? G.nx : -NG) + r % NG;
  int iy = (hy ? G.ny : -NG) + (r / NG) % NG;
  int iz = (hz ? G.nz : -NG) + r / (NG * NG);
  long s = cidx(G, hx ? G.nx - 1 : 0, hy ? G.ny - 1 : 0, hz ? G.nz - 1 : 0);
  long d = cidx(G, ix, iy, iz);
#pragma unroll
  for (int iv = 0; iv < NV; iv++) p[iv * G.ncell + d] = p[iv * G.ncell + s];
  p2u_at<MAG, RAD>(P, G, p, u, ix, iy, iz);
}

// ---------------------------------------------------------------------------
// wave speeds on loop_1 (domain +- lim) and the time-step reduction
// ---------------------------------------------------------------------------
__device__ __forceinline__ void atomic_max_pos(double *addr, double v)
{
  // non-negative doubles order as their bit patterns
  atomicMax((unsigned long long *)addr, (unsigned long long)__double_as_longlong(v));
}

template <int MAG, int RAD>
__global__ void __launch_bounds__(BS) k_wavespeeds(koral_params P, DGrid G, const double *__restrict__ p,
                                                   double *__restrict__ ws, double *tsd_max, Box b)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  __shared__ double red[BS / 64];
  long i = (long)blockIdx.x * BS + threadIdx.x;
  double td = 0.;
  if (i < b.size()) {
    int ix, iy, iz;
    b.at(i, ix, iy, iz);
    int out = (ix < 0 || ix >= G.nx) + (iy < 0 || iy >= G.ny) + (iz < 0 || iz >= G.nz);
    if (MAG || out <= 1) {
      long c = cidx(G, ix, iy, iz);
      double pp[NV], a[NWS];
      load_p<MAG, RAD>(p, G.ncell, c, pp);
      Geom g;
      geom_cell(G, ix, iy, g);
      wavespeeds_cell<MAG, RAD>(P, G, pp, ix, iy, iz, g, a);
#pragma unroll
      for (int k = 0; k < NWS; k++) ws[k * G.ncell + c] = a[k];
      if (out == 0) {
        double w[3];
#pragma unroll
        for (int d = 0; d < 3; d++) w[d] = RAD ? mx(a[12 + d], a[15 + d]) : a[12 + d];
        double dx = size_x(G, ix, 0), dy = size_x(G, iy, 1), dz = size_x(G, iz, 2);
        double t;
        if (G.nz > 1 && G.ny > 1)
          t = w[0] / dx + w[1] / dy + w[2] / dz;
        else if (G.nz == 1 && G.ny > 1)
          t = w[0] / dx + w[1] / dy;
        else if (G.ny == 1 && G.nz > 1)
          t = w[0] / dx + w[2] / dz;
        else
          t = w[0] / dx;
        t /= P.tsteplim;
        td = t;
      }
    }
  }
  // wave-level max, then block, then one atomic per block
  for (int off = 32; off > 0; off >>= 1) td = mx(td, __shfl_xor(td, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = td;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int k = 1; k < BS / 64; k++) m = mx(m, red[k]);
    if (m > 0.) atomic_max_pos(tsd_max, m);
  }
}

// ---------------------------------------------------------------------------
// per face: reconstruct both sides, face floors, fluxes, Riemann solver
// (finite.c:694-1014 and f_calc_fluxes_at_faces finite.c:1297-1415)
// ---------------------------------------------------------------------------
template <int MAG, int RAD, int ORDER>
__device__ __forceinline__ void recon_face(const DGrid &G, const double *__restrict__ p, int d, int ix, int iy,
                                           int iz, double *pL, double *pR)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  // cells -3..+2 relative to the face (face between cells -1 and 0)
  long c[6];
  double dx[6];
#pragma unroll
  for (int o = 0; o < 6; o++) {
    int q[3] = {ix, iy, iz};
    q[d] += o - 3;
    const bool need = (ORDER > 1) || (o >= 1 && o <= 4);
    c[o] = need ? cidx(G, q[0], q[1], q[2]) : 0;
    dx[o] = (ORDER == 2 && need) ? size_x(G, q[d], d) : 0.;
  }
  const long n = G.ncell;
#pragma unroll
  for (int iv = 0; iv < NV; iv++) {
    if (ORDER == 4) {
      double v0 = p[iv * n + c[0]], v1 = p[iv * n + c[1]], v2 = p[iv * n + c[2]], v3 = p[iv * n + c[3]],
             v4 = p[iv * n + c[4]], v5 = p[iv * n + c[5]];
      pL[iv] = mp5(v0, v1, v2, v3, v4);  // right face of cell -1
      pR[iv] = mp5(v5, v4, v3, v2, v1);  // left face of cell 0
    } else if (ORDER == 2) {
      double v[6];
#pragma unroll
      for (int o = 0; o < 6; o++) v[o] = p[iv * n + c[o]];
      double ul, ur;
      ppm(v, dx, ul, ur);
      pL[iv] = ur;
      ppm(v + 1, dx + 1, ul, ur);
      pR[iv] = ul;
    }
  }
}

template <int MAG, int RAD, int ORDER>
__global__ void __launch_bounds__(BS) k_flux(koral_params P, DGrid G, const double *__restrict__ p,
                                             const double *__restrict__ ws, double *__restrict__ flb, int d, Box b)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  constexpr int NVMHD = 6 + 3 * MAG;
  long i = (long)blockIdx.x * BS + threadIdx.x;
  if (i >= b.size()) return;
  int ix, iy, iz;
  b.at(i, ix, iy, iz);
  double pL[NV], pR[NV];
  if (ORDER == 1) {
    long c[4];
    const long n = G.ncell;
#pragma unroll
    for (int o = 0; o < 4; o++) {
      int q[3] = {ix, iy, iz};
      q[d] += o - 2;
      c[o] = cidx(G, q[0], q[1], q[2]);
    }
    const double th = P.minmod_theta;
#pragma unroll
    for (int iv = 0; iv < NV; iv++) {
      double um1 = p[iv * n + c[0]], u0 = p[iv * n + c[1]], up1 = p[iv * n + c[2]];
      pL[iv] = u0 + .5 * minmod3(th * (up1 - u0), .5 * (up1 - um1), th * (u0 - um1));
      double vm1 = u0, v0 = up1, vp1 = p[iv * n + c[3]];
      pR[iv] = v0 - .5 * minmod3(th * (vp1 - v0), .5 * (vp1 - vm1), th * (v0 - vm1));
    }
  } else {
    recon_face<MAG, RAD, ORDER>(G, p, d, ix, iy, iz, pL, pR);
  }
  Geom g;
  geom_face(G, d, ix, iy, g);
  floors_mhd<MAG, RAD>(P, pL, g);
  floors_mhd<MAG, RAD>(P, pR, g);
  double fL[NV], fR[NV], uL[NV], uR[NV];
  flux_prime<MAG, RAD>(P, pL, d + 1, g, fL);
  flux_prime<MAG, RAD>(P, pR, d + 1, g, fR);
  kd::p2u<MAG, RAD>(P, pL, uL, g);
  kd::p2u<MAG, RAD>(P, pR, uR, g);
  int qm[3] = {ix, iy, iz};
  qm[d] -= 1;
  const long cR = cidx(G, ix, iy, iz), cL = cidx(G, qm[0], qm[1], qm[2]);
  const long n = G.ncell;
  const long f = fidx(G, d, ix, iy, iz);
  const long nf = d == 0 ? G.nfx : (d == 1 ? G.nfy : G.nfz);
#pragma unroll
  for (int iv = 0; iv < NV; iv++) {
    const int r = (RAD && iv >= NVMHD) ? 1 : 0;
    double ag = mx(ws[(12 + 3 * r + d) * n + cR], ws[(12 + 3 * r + d) * n + cL]);
    double al = mn(ws[(6 * r + 2 * d) * n + cR], ws[(6 * r + 2 * d) * n + cL]);
    double ar = mx(ws[(6 * r + 2 * d + 1) * n + cR], ws[(6 * r + 2 * d + 1) * n + cL]);
    double F;
    if (P.fluxmethod == 0) {
      F = .5 * (fR[iv] + fL[iv] - ag * (uR[iv] - uL[iv]));
    } else {
      if (al > 0.)
        F = fL[iv];
      else if (ar < 0.)
        F = fL[iv];
      else
        F = (-al * fR[iv] + ar * fL[iv] + al * ar * (uR[iv] - uL[iv])) / (ar - al);
    }
    flb[iv * nf + f] = F;
  }
}

// ---------------------------------------------------------------------------
// flux-CT (magn.c:199-339)
// ---------------------------------------------------------------------------
struct Flx {
  double *x, *y, *z;
  long nfx, nfy, nfz;
};

template <int NV>
__global__ void __launch_bounds__(BS) k_emf(DGrid G, Flx F, double *emf)
{
  const int NX = G.nx, NY = G.ny, NZ = G.nz;
  long t = (long)blockIdx.x * BS + threadIdx.x;
  long ne = (long)(NX + 1) * (NY + 1) * (NZ + 1);
  if (t >= ne) return;
  int ix = (int)(t % (NX + 1));
  int iy = (int)((t / (NX + 1)) % (NY + 1));
  int iz = (int)(t / ((long)(NX + 1) * (NY + 1)));
  const double c1 = (NY > 1 && NZ > 1) ? 0.25 : 0.5;
  const double c2 = (NX > 1 && NZ > 1) ? 0.25 : 0.5;
  const double c3 = (NX > 1 && NY > 1) ? 0.25 : 0.5;
  auto FX = [&](int iv, int a, int b, int c) { return F.x[iv * F.nfx + fxidx(G, a, NY > 1 ? b : 0, NZ > 1 ? c : 0)]; };
  auto FY = [&](int iv, int a, int b, int c) { return F.y[iv * F.nfy + fyidx(G, NX > 1 ? a : 0, b, NZ > 1 ? c : 0)]; };
  auto FZ = [&](int iv, int a, int b, int c) { return F.z[iv * F.nfz + fzidx(G, NX > 1 ? a : 0, NY > 1 ? b : 0, c)]; };
  double e1 = 0., e2 = 0., e3 = 0.;
  if (NY > 1 || NZ > 1) {
    double s = 0.;
    if (NY > 1) s = s + FY(8, ix, iy, iz) + FY(8, ix, iy, iz - 1);
    if (NZ > 1) s = s - FZ(7, ix, iy, iz) - FZ(7, ix, iy - 1, iz);
    e1 = c1 * (s);
  }
  if (NX > 1 || NZ > 1) {
    double s = 0.;
    if (NZ > 1) s = s + FZ(6, ix, iy, iz) + FZ(6, ix - 1, iy, iz);
    if (NX > 1) s = s - FX(8, ix, iy, iz) - FX(8, ix, iy, iz - 1);
    e2 = c2 * (s);
  }
  if (NX > 1 || NY > 1) {
    double s = 0.;
    if (NX > 1) s = s + FX(7, ix, iy, iz) + FX(7, ix, iy - 1, iz);
    if (NY > 1) s = s - FY(6, ix, iy, iz) - FY(6, ix - 1, iy, iz);
    e3 = c3 * (s);
  }
  emf[0 * ne + t] = e1;
  emf[1 * ne + t] = e2;
  emf[2 * ne + t] = e3;
}

__global__ void __launch_bounds__(BS) k_ct_faces(DGrid G, Flx F, const double *emf)
{
  const int NX = G.nx, NY = G.ny, NZ = G.nz;
  long t = (long)blockIdx.x * BS + threadIdx.x;
  long ne = (long)(NX + 1) * (NY + 1) * (NZ + 1);
  if (t >= ne) return;
  int ix = (int)(t % (NX + 1));
  int iy = (int)((t / (NX + 1)) % (NY + 1));
  int iz = (int)(t / ((long)(NX + 1) * (NY + 1)));
  auto E = [&](int k, int a, int b, int c) { return emf[(k - 1) * ne + a + (long)b * (NX + 1) + (long)c * (NY + 1) * (NX + 1)]; };
  if (NX > 1 && iy < NY && iz < NZ) {
    long f = fxidx(G, ix, iy, iz);
    F.x[6 * F.nfx + f] = 0.;
    F.x[7 * F.nfx + f] = 0.5 * (E(3, ix, iy, iz) + E(3, ix, iy + 1, iz));
    F.x[8 * F.nfx + f] = -0.5 * (E(2, ix, iy, iz) + E(2, ix, iy, iz + 1));
  }
  if (NY > 1 && ix < NX && iz < NZ) {
    long f = fyidx(G, ix, iy, iz);
    F.y[6 * F.nfy + f] = -0.5 * (E(3, ix, iy, iz) + E(3, ix + 1, iy, iz));
    F.y[7 * F.nfy + f] = 0.;
    F.y[8 * F.nfy + f] = 0.5 * (E(1, ix, iy, iz) + E(1, ix, iy, iz + 1));
  }
  if (NZ > 1 && ix < NX && iy < NY) {
    long f = fzidx(G, ix, iy, iz);
    F.z[6 * F.nfz + f] = 0.5 * (E(2, ix, iy, iz) + E(2, ix + 1, iy, iz));
    F.z[7 * F.nfz + f] = -0.5 * (E(1, ix, iy, iz) + E(1, ix, iy + 1, iz));
    F.z[8 * F.nfz + f] = 0.;
  }
}

// ---------------------------------------------------------------------------
// conservative update with metric sources (finite.c:1062-1160)
// ---------------------------------------------------------------------------
template <int MAG, int RAD>
__global__ void __launch_bounds__(BS) k_update(koral_params P, DGrid G, const double *__restrict__ p,
                                               double *__restrict__ u, Flx F, double dt)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  long i = (long)blockIdx.x * BS + threadIdx.x;
  Box b{0, 0, 0, G.nx, G.ny, G.nz};
  if (i >= b.size()) return;
  int ix, iy, iz;
  b.at(i, ix, iy, iz);
  long c = cidx(G, ix, iy, iz);
  double pp[NV], ms[NV];
  load_p<MAG, RAD>(p, G.ncell, c, pp);
  Geom g;
  geom_cell(G, ix, iy, g);
  metric_source<MAG, RAD>(P, G, pp, ix, iy, g, ms);
  double dx = size_x(G, ix, 0), dy = size_x(G, iy, 1), dz = size_x(G, iz, 2);
  long fxl = fxidx(G, ix, iy, iz), fxr = fxidx(G, ix + 1, iy, iz);
  long fyl = 0, fyr = 0, fzl = 0, fzr = 0;
  if (G.ny > 1) { fyl = fyidx(G, ix, iy, iz); fyr = fyidx(G, ix, iy + 1, iz); }
  if (G.nz > 1) { fzl = fzidx(G, ix, iy, iz); fzr = fzidx(G, ix, iy, iz + 1); }
#pragma unroll
  for (int iv = 0; iv < NV; iv++) {
    double du = -(F.x[iv * F.nfx + fxr] - F.x[iv * F.nfx + fxl]) * dt / dx;
    if (G.ny > 1) du = du - (F.y[iv * F.nfy + fyr] - F.y[iv * F.nfy + fyl]) * dt / dy;
    if (G.nz > 1) du = du - (F.z[iv * F.nfz + fzr] - F.z[iv * F.nfz + fzl]) * dt / dz;
    u[iv * G.ncell + c] = u[iv * G.ncell + c] + du + ms[iv] * dt;
  }
}

// ---------------------------------------------------------------------------
// u2p sweep (calc_u2p finite.c:502-561 / calc_primitives u2p.c:13-133)
// ---------------------------------------------------------------------------
template <int MAG, int RAD>
__global__ void __launch_bounds__(BS) k_u2p_grid(koral_params P, DGrid G, double *__restrict__ u,
                                                 double *__restrict__ p, signed char *__restrict__ fl,
                                                 int *__restrict__ counters)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  long i = (long)blockIdx.x * BS + threadIdx.x;
  Box b{0, 0, 0, G.nx, G.ny, G.nz};
  int nhd = 0, nrad = 0, nent = 0;
  if (i < b.size()) {
    int ix, iy, iz;
    b.at(i, ix, iy, iz);
    long c = cidx(G, ix, iy, iz);
    double pp[NV], uu[NV];
    load_p<MAG, RAD>(p, G.ncell, c, pp);
    load_p<MAG, RAD>(u, G.ncell, c, uu);
    Geom g;
    geom_cell(G, ix, iy, g);
    int fe, fe2, fh, fr;
    calc_primitives<MAG, RAD>(P, uu, pp, g, fe, fe2, fh, fr);
    store_p<MAG, RAD>(p, G.ncell, c, pp);
    store_p<MAG, RAD>(u, G.ncell, c, uu);
    const long n = G.ncell;
    fl[F_ENTROPY * n + c] = (signed char)fe;
    fl[F_ENTROPY2 * n + c] = (signed char)fe2;
    fl[F_HDFIXUP * n + c] = (signed char)fh;
    fl[F_RADFIXUP * n + c] = (signed char)fr;
    nhd = fh != 0;
    nrad = fr != 0;
    nent = fe != 0;
  }
  // wave-level counts
  unsigned long long mh = __ballot(nhd), mr = __ballot(nrad), me = __ballot(nent);
  if ((threadIdx.x & 63) == 0) {
    if (mh) atomicAdd(&counters[3], __popcll(mh));
    if (mr) atomicAdd(&counters[4], __popcll(mr));
    if (me) atomicAdd(&counters[5], __popcll(me));
  }
}

// ---------------------------------------------------------------------------
// cell_fixup (finite.c:4468-4640): cells are processed in (ix, iy, iz) order
// in the reference and clear their flag when processed, so a lower neighbour
// always counts as good and an upper neighbour only if it was not flagged.
// Results go to scratch (pn, un) and are applied afterwards, so every cell
// reads the neighbours' values from before the sweep.
// ---------------------------------------------------------------------------
template <int MAG, int RAD>
__global__ void __launch_bounds__(BS) k_fixup(koral_params P, DGrid G, const double *__restrict__ p,
                                              signed char *__restrict__ fl, int type, double *__restrict__ pn,
                                              double *__restrict__ un, int *__restrict__ mark)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  constexpr int EE = Lay<MAG, RAD>::EE;
  long i = (long)blockIdx.x * BS + threadIdx.x;
  Box b{0, 0, 0, G.nx, G.ny, G.nz};
  if (i >= b.size()) return;
  int ix, iy, iz;
  b.at(i, ix, iy, iz);
  const long n = G.ncell;
  long c = cidx(G, ix, iy, iz);
  mark[c] = 0;
  if (fl[type * n + c] == 0) return;
  const int NX = G.nx, NY = G.ny, NZ = G.nz;
  const int nb[6][3] = {{ix - 1, iy, iz}, {ix + 1, iy, iz}, {ix, iy - 1, iz},
                        {ix, iy + 1, iz}, {ix, iy, iz - 1}, {ix, iy, iz + 1}};
  const bool inb[6] = {ix - 1 >= 0, ix + 1 < NX, iy - 1 >= 0, iy + 1 < NY, iz - 1 >= 0, iz + 1 < NZ};
  long cn[6];
  int in = 0;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    if (!inb[k]) continue;
    long q = cidx(G, nb[k][0], nb[k][1], nb[k][2]);
    bool lower = (k & 1) == 0;
    if (lower || fl[type * n + q] == 0) cn[in++] = q;
  }
  if ((NZ == 1 && NY == 1 && in >= 1) || (NZ == 1 && in >= 2) || (NY == 1 && in >= 2) || in >= 3) {
    double pp[NV], uu[NV];
#pragma unroll
    for (int iv = 0; iv < NV; iv++) {
      bool fix = false;
      if (type == F_HDFIXUP && iv != 0 && iv < 6) fix = true;
      if (type == F_RADFIXUP && iv != 0 && RAD && iv >= EE && iv <= EE + 3) fix = true;
      if (type == F_RADIMPFIXUP && iv != 0 && (iv < 6 || (RAD && iv >= EE && iv <= EE + 3))) fix = true;
      if (fix) {
        double s = 0.;
        for (int q = 0; q < in; q++) s += p[iv * n + cn[q]];
        pp[iv] = s / (double)in;
      } else {
        pp[iv] = p[iv * n + c];
      }
    }
    Geom g;
    geom_cell(G, ix, iy, g);
    kd::p2u<MAG, RAD>(P, pp, uu, g);
#pragma unroll
    for (int iv = 0; iv < 6 + 3 * MAG; iv++) {
      pn[iv * n + c] = pp[iv];
      un[iv * n + c] = uu[iv];
    }
    mark[c] = 1;
  }
}

template <int MAG, int RAD>
__global__ void __launch_bounds__(BS) k_fixup_apply(DGrid G, double *__restrict__ p, double *__restrict__ u,
                                                    signed char *__restrict__ fl, int type,
                                                    const double *__restrict__ pn, const double *__restrict__ un,
                                                    const int *__restrict__ mark)
{
  long i = (long)blockIdx.x * BS + threadIdx.x;
  Box b{0, 0, 0, G.nx, G.ny, G.nz};
  if (i >= b.size()) return;
  int ix, iy, iz;
  b.at(i, ix, iy, iz);
  const long n = G.ncell;
  long c = cidx(G, ix, iy, iz);
  fl[type * n + c] = 0;
  if (!mark[c]) return;
#pragma unroll
  for (int iv = 0; iv < 6 + 3 * MAG; iv++) {
    p[iv * n + c] = pn[iv * n + c];
    u[iv * n + c] = un[iv * n + c];
  }
}

// ---------------------------------------------------------------------------
// implicit sweep (op_implicit finite.c:1242-1292)
// ---------------------------------------------------------------------------
template <int MAG, int RAD>
__global__ void __launch_bounds__(BS) k_implicit_grid(koral_params P, DGrid G, double *__restrict__ u,
                                                      double *__restrict__ p, signed char *__restrict__ fl, double dt,
                                                      double t, int *__restrict__ counters)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  long i = (long)blockIdx.x * BS + threadIdx.x;
  Box b{0, 0, 0, G.nx, G.ny, G.nz};
  int ok = 0, bad = 0, its = 0;
  if (i < b.size()) {
    int ix, iy, iz;
    b.at(i, ix, iy, iz);
    long c = cidx(G, ix, iy, iz);
    double pp[NV], uu[NV];
    load_p<MAG, RAD>(p, G.ncell, c, pp);
    load_p<MAG, RAD>(u, G.ncell, c, uu);
    Geom g;
    geom_cell(G, ix, iy, g);
    ImpInfo inf;
    int r = solve_implicit_lab<MAG, RAD>(P, uu, pp, g, dt, t, inf);
    if (r == 0) {
      store_p<MAG, RAD>(p, G.ncell, c, pp);
      store_p<MAG, RAD>(u, G.ncell, c, uu);
    }
    const long n = G.ncell;
    fl[F_RADSOURCETYPE * n + c] = 3;
    fl[F_RADFIXUP * n + c] = r == 0 ? 0 : -1;
    fl[F_RADIMPFIXUP * n + c] = r == 0 ? 0 : -1;
    ok = r == 0;
    bad = r != 0;
    its = inf.iters;
  }
  unsigned long long mo = __ballot(ok), mb = __ballot(bad);
  for (int off = 32; off > 0; off >>= 1) its += __shfl_xor(its, off);
  if ((threadIdx.x & 63) == 0) {
    if (mo) atomicAdd(&counters[0], __popcll(mo));
    if (its) atomicAdd(&counters[1], its);
    if (mb) atomicAdd(&counters[2], __popcll(mb));
  }
}

// ---------------------------------------------------------------------------
// RK2IMEX linear combinations over the domain (ko.c:444-570)
// out = a*x + b*y (+ c*z)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BS) k_comb(DGrid G, int nv, double *out, double a, const double *x, double b,
                                             const double *y, double c, const double *z)
{
  long i = (long)blockIdx.x * BS + threadIdx.x;
  long nd = (long)G.nx * G.ny * G.nz;
  if (i >= nd * nv) return;
  int iv = (int)(i / nd);
  long r = i % nd;
  int ix = (int)(r % G.nx), iy = (int)((r / G.nx) % G.ny), iz = (int)(r / ((long)G.nx * G.ny));
  long k = (long)iv * G.ncell + cidx(G, ix, iy, iz);
  double v;
  if (z)
    v = a * x[k] + b * y[k] + c * z[k];
  else if (y)
    v = a * x[k] + b * y[k];
  else
    v = a * x[k];
  out[k] = v;
}

// AoS (KORAL host layout, [cell][nv]) <-> SoA
__global__ void k_aos2soa(long ncell, int nv, const double *aos, double *soa)
{
  long i = (long)blockIdx.x * BS + threadIdx.x;
  if (i >= ncell * nv) return;
  long c = i / nv;
  int iv = (int)(i % nv);
  soa[iv * ncell + c] = aos[i];
}
__global__ void k_soa2aos(long ncell, int nv, const double *soa, double *aos)
{
  long i = (long)blockIdx.x * BS + threadIdx.x;
  if (i >= ncell * nv) return;
  long c = i / nv;
  int iv = (int)(i % nv);
  aos[i] = soa[iv * ncell + c];
}
__global__ void k_flags_i2c(long ncell, const int *aos, signed char *soa)
{
  long i = (long)blockIdx.x * BS + threadIdx.x;
  if (i >= ncell * NFLAGS) return;
  long c = i / NFLAGS;
  int f = (int)(i % NFLAGS);
  soa[f * ncell + c] = (signed char)aos[i];
}
__global__ void k_flags_c2i(long ncell, const signed char *soa, int *aos)
{
  long i = (long)blockIdx.x * BS + threadIdx.x;
  if (i >= ncell * NFLAGS) return;
  long c = i / NFLAGS;
  int f = (int)(i % NFLAGS);
  aos[i] = (int)soa[f * ncell + c];
}

}  // namespace

// ---------------------------------------------------------------------------
// host-side grid object
// ---------------------------------------------------------------------------
struct koral_hip_grid {
  koral_params P;
  DGrid G;
  hipStream_t s = nullptr;
  std::vector<void *> allocs;
  double *p = nullptr, *u = nullptr;
  signed char *fl = nullptr;
  double *ws = nullptr, *flb[3] = {nullptr, nullptr, nullptr}, *emf = nullptr;
  double *pn = nullptr, *un = nullptr;
  int *mark = nullptr;
  double *ut0, *ut1, *uforget, *dut1, *dut2, *drt1, *drt2;
  double *tsd = nullptr;   // device tstepdenmax
  int *counters = nullptr; // device counters[8]
  double *stage = nullptr; // AoS staging
  long ncell = 0;

  template <class T>
  int alloc(T **ptr, size_t count)
  {
    void *q = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess) return check_hip(e, "hipMalloc");
    allocs.push_back(q);
    *ptr = (T *)q;
    return 0;
  }
  ~koral_hip_grid()
  {
    for (void *q : allocs) (void)hipFree(q);
  }
};

#define DISPATCH(P, KERNEL, ...)                                               \
  do {                                                                         \
    if ((P).mag && (P).rad) KERNEL<1, 1> __VA_ARGS__;                          \
    else if ((P).mag) KERNEL<1, 0> __VA_ARGS__;                                \
    else if ((P).rad) KERNEL<0, 1> __VA_ARGS__;                                \
    else KERNEL<0, 0> __VA_ARGS__;                                             \
  } while (0)

namespace {

inline int dimn_host(const DGrid &G, int d) { return d == 0 ? G.nx : (d == 1 ? G.ny : G.nz); }

int run_set_bc(koral_hip_grid *g)
{
  const DGrid &G = g->G;
  Box b{-G.ngx, -G.ngy, -G.ngz, G.sx, G.sy, G.sz};
  DISPATCH(g->P, k_bc_faces, <<<blocks_for(b.size()), BS, 0, g->s>>>(g->P, G, g->p, g->u, b));
  if (G.ny > 1) {
    const int ng = G.ng;
    auto edge = [&](int a, int bb) {
      int c = 3 - a - bb;
      int ngc = c == 0 ? G.ngx : (c == 1 ? G.ngy : G.ngz);
      long nt = 4L * (dimn_host(G, c) + 2 * ngc);
      DISPATCH(g->P, k_bc_edge, <<<blocks_for(nt), BS, 0, g->s>>>(g->P, G, g->p, g->u, a, bb));
    };
    edge(0, 1);
    if (G.nz > 1) {
      edge(0, 2);
      edge(1, 2);
      long nt = 8L * ng * ng * ng;
      DISPATCH(g->P, k_bc_total, <<<blocks_for(nt), BS, 0, g->s>>>(g->P, G, g->p, g->u));
    }
  }
  return check_launch("set_bc");
}

}  // namespace

namespace {

int run_explicit(koral_hip_grid *g, double t, double dt)
{
  (void)t;
  const DGrid &G = g->G;
  const koral_params &P = g->P;
  int rc = run_set_bc(g);
  if (rc) return rc;
  int lim = (P.int_order == 4) ? 2 : 1;
  int lx = lim, ly = G.ny > 1 ? lim : 0, lz = G.nz > 1 ? lim : 0;
  Box wb{-lx, -ly, -lz, G.nx + 2 * lx, G.ny + 2 * ly, G.nz + 2 * lz};
  DISPATCH(P, k_wavespeeds, <<<blocks_for(wb.size()), BS, 0, g->s>>>(P, G, g->p, g->ws, g->tsd, wb));
  int tt = P.mag ? 1 : 0;
  int n[3] = {G.nx, G.ny, G.nz};
  for (int d = 0; d < 3; d++) {
    if (n[d] == 1) continue;
    int lo[3], cnt[3];
    for (int k = 0; k < 3; k++) {
      if (k == d) { lo[k] = 0; cnt[k] = n[k] + 1; }
      else if (n[k] == 1) { lo[k] = 0; cnt[k] = 1; }
      else { lo[k] = -tt; cnt[k] = n[k] + 2 * tt; }
    }
    Box fb{lo[0], lo[1], lo[2], cnt[0], cnt[1], cnt[2]};
    auto launch = [&](auto kern) {
      kern<<<blocks_for(fb.size()), BS, 0, g->s>>>(P, G, g->p, g->ws, g->flb[d], d, fb);
    };
    if (P.mag && P.rad) {
      if (P.int_order == 4) launch(k_flux<1, 1, 4>); else if (P.int_order == 2) launch(k_flux<1, 1, 2>); else launch(k_flux<1, 1, 1>);
    } else if (P.mag) {
      if (P.int_order == 4) launch(k_flux<1, 0, 4>); else if (P.int_order == 2) launch(k_flux<1, 0, 2>); else launch(k_flux<1, 0, 1>);
    } else if (P.rad) {
      if (P.int_order == 4) launch(k_flux<0, 1, 4>); else if (P.int_order == 2) launch(k_flux<0, 1, 2>); else launch(k_flux<0, 1, 1>);
    } else {
      if (P.int_order == 4) launch(k_flux<0, 0, 4>); else if (P.int_order == 2) launch(k_flux<0, 0, 2>); else launch(k_flux<0, 0, 1>);
    }
  }
  Flx F{g->flb[0], g->flb[1], g->flb[2], G.nfx, G.nfy, G.nfz};
  if (P.mag) {
    long ne = (long)(G.nx + 1) * (G.ny + 1) * (G.nz + 1);
    k_emf<9><<<blocks_for(ne), BS, 0, g->s>>>(G, F, g->emf);
    k_ct_faces<<<blocks_for(ne), BS, 0, g->s>>>(G, F, g->emf);
  }
  long nd = (long)G.nx * G.ny * G.nz;
  DISPATCH(P, k_update, <<<blocks_for(nd), BS, 0, g->s>>>(P, G, g->p, g->u, F, dt));
  return check_launch("op_explicit");
}

int run_fixup(koral_hip_grid *g, int type)
{
  const koral_params &P = g->P;
  if (!P.dofixups) return 0;
  if (type == F_HDFIXUP && !P.dou2pmhdfixups) return 0;
  if (type == F_RADFIXUP && !P.dou2pradfixups) return 0;
  if (type == F_RADIMPFIXUP && !P.doradimpfixups) return 0;
  const DGrid &G = g->G;
  long nd = (long)G.nx * G.ny * G.nz;
  DISPATCH(P, k_fixup, <<<blocks_for(nd), BS, 0, g->s>>>(P, G, g->p, g->fl, type, g->pn, g->un, g->mark));
  DISPATCH(P, k_fixup_apply, <<<blocks_for(nd), BS, 0, g->s>>>(G, g->p, g->u, g->fl, type, g->pn, g->un, g->mark));
  return check_launch("cell_fixup");
}

int run_u2p(koral_hip_grid *g)
{
  const DGrid &G = g->G;
  long nd = (long)G.nx * G.ny * G.nz;
  DISPATCH(g->P, k_u2p_grid, <<<blocks_for(nd), BS, 0, g->s>>>(g->P, G, g->u, g->p, g->fl, g->counters));
  int rc = run_fixup(g, F_HDFIXUP);
  if (!rc && g->P.rad) rc = run_fixup(g, F_RADFIXUP);
  if (!rc) rc = run_set_bc(g);
  return rc ? rc : check_launch("calc_u2p");
}

int run_implicit(koral_hip_grid *g, double t, double dt)
{
  if (!g->P.rad) return 0;
  const DGrid &G = g->G;
  long nd = (long)G.nx * G.ny * G.nz;
  if (g->P.mag)
    k_implicit_grid<1, 1><<<blocks_for(nd), BS, 0, g->s>>>(g->P, G, g->u, g->p, g->fl, dt, t, g->counters);
  else
    k_implicit_grid<0, 1><<<blocks_for(nd), BS, 0, g->s>>>(g->P, G, g->u, g->p, g->fl, dt, t, g->counters);
  int rc = check_launch("op_implicit");
  return rc ? rc : run_fixup(g, F_RADIMPFIXUP);
}

void comb(koral_hip_grid *g, double *out, double a, const double *x, double b = 0., const double *y = nullptr,
          double c = 0., const double *z = nullptr)
{
  const DGrid &G = g->G;
  long nd = (long)G.nx * G.ny * G.nz * g->P.nv;
  k_comb<<<blocks_for(nd), BS, 0, g->s>>>(G, g->P.nv, out, a, x, b, y, c, z);
}

int run_step(koral_hip_grid *g, double t, double dt)
{
  const double gamma = 1. - 1. / sqrt(2.);
  int rc = 0;
  comb(g, g->ut0, 1., g->u);
  if ((rc = run_implicit(g, t, dt * gamma))) return rc;
  comb(g, g->drt1, (1. / (dt * gamma)), g->u, (-1. / (dt * gamma)), g->ut0);
  comb(g, g->ut1, 1., g->u);
  if ((rc = run_u2p(g))) return rc;
  if ((rc = run_explicit(g, t, dt))) return rc;
  comb(g, g->dut1, (1. / (dt)), g->u, (-1. / (dt)), g->ut1);
  comb(g, g->u, 1., g->ut0, (dt), g->dut1, (dt * (1. - 2. * gamma)), g->drt1);
  comb(g, g->uforget, 1., g->u);
  if ((rc = run_u2p(g))) return rc;
  if ((rc = run_implicit(g, t, gamma * dt))) return rc;
  comb(g, g->drt2, (1. / (dt * gamma)), g->u, (-1. / (dt * gamma)), g->uforget);
  comb(g, g->ut1, 1., g->u);  // ut2
  if ((rc = run_u2p(g))) return rc;
  if ((rc = run_explicit(g, t, dt))) return rc;
  comb(g, g->dut2, (1. / (dt)), g->u, (-1. / (dt)), g->ut1);
  comb(g, g->u, 1., g->ut0, (dt / 2.), g->dut1, (dt / 2.), g->dut2);
  comb(g, g->u, 1., g->u, (dt / 2.), g->drt1, (dt / 2.), g->drt2);
  return run_u2p(g);
}

}  // namespace

extern "C" {

int koral_hip_grid_create(const koral_params *P, const koral_grid_desc *D, koral_hip_grid **out)
{
  if (check_params(P) || !D || !out) return set_error(KORAL_E_ARG, "grid_create: bad arguments");
  if (D->nx < 1 || D->ny < 1 || D->nz < 1 || D->ng < 2 || D->ng > 4)
    return set_error(KORAL_E_ARG, "grid_create: bad sizes");
  if ((P->int_order == 4 && D->ng < 4) || (P->int_order == 2 && D->ng < 3))
    return set_error(KORAL_E_ARG, "grid_create: NG too small for INT_ORDER");
  if (D->ny == 1 && D->nz > 1) return set_error(KORAL_E_ARG, "grid_create: NY=1 with NZ>1 not supported");
  auto *g = new koral_hip_grid();
  g->P = *P;
  DGrid &G = g->G;
  G.nx = D->nx; G.ny = D->ny; G.nz = D->nz; G.ng = D->ng;
  G.ngx = D->ng; G.ngy = D->ny > 1 ? D->ng : 0; G.ngz = D->nz > 1 ? D->ng : 0;
  G.sx = G.nx + 2 * G.ngx; G.sy = G.ny + 2 * G.ngy; G.sz = G.nz + 2 * G.ngz;
  G.ncell = (long)G.sx * G.sy * G.sz;
  G.nfx = (long)(G.sx + 1) * G.sy * G.sz;
  G.nfy = (long)G.sx * (G.sy + 1) * G.sz;
  G.nfz = (long)G.sx * G.sy * (G.sz + 1);
  G.npl = (long)G.sy * G.sx;
  G.nplx = (long)G.sy * (G.sx + 1);
  G.nply = (long)(G.sy + 1) * G.sx;
  for (int i = 0; i < 6; i++) G.bc[i] = D->bc[i];
  g->ncell = G.ncell;
  const int nv = P->nv;
  int rc = 0;
  auto up = [&](const double *h, size_t n, const double **dst) {
    if (rc) return;
    double *d = nullptr;
    rc = g->alloc(&d, n);
    if (!rc) rc = check_hip(hipMemcpy(d, h, n * sizeof(double), hipMemcpyHostToDevice), "upload table");
    *dst = d;
  };
  up(D->xb[0], D->nx + 2 * D->ng + 1, &G.xb0);
  up(D->xb[1], D->ny + 2 * D->ng + 1, &G.xb1);
  up(D->xb[2], D->nz + 2 * D->ng + 1, &G.xb2);
  up(D->gc, 20 * G.npl, &G.gc);
  up(D->Gc, 20 * G.npl, &G.Gc);
  up(D->kr, 64 * G.npl, &G.kr);
  up(D->gfx, 20 * G.nplx, &G.gfx);
  up(D->Gfx, 20 * G.nplx, &G.Gfx);
  up(D->gfy, 20 * G.nply, &G.gfy);
  up(D->Gfy, 20 * G.nply, &G.Gfy);
  long nc = G.ncell;
  if (!rc) rc = g->alloc(&g->p, nv * nc);
  if (!rc) rc = g->alloc(&g->u, nv * nc);
  if (!rc) rc = g->alloc(&g->fl, NFLAGS * nc);
  if (!rc) rc = g->alloc(&g->ws, NWS * nc);
  if (!rc && G.nx > 1) rc = g->alloc(&g->flb[0], nv * G.nfx);
  if (!rc && G.ny > 1) rc = g->alloc(&g->flb[1], nv * G.nfy);
  if (!rc && G.nz > 1) rc = g->alloc(&g->flb[2], nv * G.nfz);
  if (!rc) rc = g->alloc(&g->emf, 3L * (G.nx + 1) * (G.ny + 1) * (G.nz + 1));
  if (!rc) rc = g->alloc(&g->pn, nv * nc);
  if (!rc) rc = g->alloc(&g->un, nv * nc);
  if (!rc) rc = g->alloc(&g->mark, nc);
  double **work[7] = {&g->ut0, &g->ut1, &g->uforget, &g->dut1, &g->dut2, &g->drt1, &g->drt2};
  for (int k = 0; k < 7 && !rc; k++) rc = g->alloc(work[k], nv * nc);
  if (!rc) rc = g->alloc(&g->tsd, 1);
  if (!rc) rc = g->alloc(&g->counters, 8);
  if (!rc) rc = g->alloc(&g->stage, nv * nc);
  if (!rc) rc = check_hip(hipMemset(g->p, 0, nv * nc * sizeof(double)), "memset");
  if (!rc) rc = check_hip(hipMemset(g->u, 0, nv * nc * sizeof(double)), "memset");
  if (!rc) rc = check_hip(hipMemset(g->fl, 0, NFLAGS * nc), "memset");
  for (int d = 0; d < 3 && !rc; d++)
    if (g->flb[d]) rc = check_hip(hipMemset(g->flb[d], 0, nv * (d == 0 ? G.nfx : d == 1 ? G.nfy : G.nfz) * sizeof(double)), "memset");
  if (!rc) rc = check_hip(hipMemset(g->ws, 0, NWS * nc * sizeof(double)), "memset");
  if (rc) {
    delete g;
    return rc;
  }
  *out = g;
  return 0;
}

int koral_hip_grid_destroy(koral_hip_grid *g)
{
  if (!g) return 0;
  (void)hipStreamSynchronize(g->s);
  delete g;
  return 0;
}

int koral_hip_grid_set_stream(koral_hip_grid *g, void *stream)
{
  if (!g) return set_error(KORAL_E_ARG, "null grid");
  g->s = (hipStream_t)stream;
  return 0;
}

int koral_hip_grid_upload(koral_hip_grid *g, const double *p_aos, const double *u_aos, const int *flags_aos)
{
  if (!g || !p_aos) return set_error(KORAL_E_ARG, "grid_upload: bad arguments");
  const long nc = g->ncell;
  const int nv = g->P.nv;
  int rc = check_hip(hipMemcpyAsync(g->stage, p_aos, nv * nc * sizeof(double), hipMemcpyHostToDevice, g->s), "upload p");
  if (rc) return rc;
  k_aos2soa<<<blocks_for(nc * nv), BS, 0, g->s>>>(nc, nv, g->stage, g->p);
  if (u_aos) {
    rc = check_hip(hipStreamSynchronize(g->s), "sync");
    if (!rc) rc = check_hip(hipMemcpyAsync(g->stage, u_aos, nv * nc * sizeof(double), hipMemcpyHostToDevice, g->s), "upload u");
    if (rc) return rc;
    k_aos2soa<<<blocks_for(nc * nv), BS, 0, g->s>>>(nc, nv, g->stage, g->u);
  }
  if (flags_aos) {
    rc = check_hip(hipStreamSynchronize(g->s), "sync");
    int *tmp = (int *)g->pn;  // scratch is large enough (nv*nc doubles >= 7*nc ints)
    if (!rc) rc = check_hip(hipMemcpyAsync(tmp, flags_aos, NFLAGS * nc * sizeof(int), hipMemcpyHostToDevice, g->s), "upload flags");
    if (rc) return rc;
    k_flags_i2c<<<blocks_for(nc * NFLAGS), BS, 0, g->s>>>(nc, tmp, g->fl);
  }
  rc = check_hip(hipStreamSynchronize(g->s), "upload sync");
  return rc ? rc : check_launch("grid_upload");
}

int koral_hip_grid_download(koral_hip_grid *g, double *p_aos, double *u_aos, int *flags_aos)
{
  if (!g) return set_error(KORAL_E_ARG, "null grid");
  const long nc = g->ncell;
  const int nv = g->P.nv;
  int rc = 0;
  if (p_aos) {
    k_soa2aos<<<blocks_for(nc * nv), BS, 0, g->s>>>(nc, nv, g->p, g->stage);
    rc = check_hip(hipMemcpyAsync(p_aos, g->stage, nv * nc * sizeof(double), hipMemcpyDeviceToHost, g->s), "download p");
    if (!rc) rc = check_hip(hipStreamSynchronize(g->s), "sync");
  }
  if (!rc && u_aos) {
    k_soa2aos<<<blocks_for(nc * nv), BS, 0, g->s>>>(nc, nv, g->u, g->stage);
    rc = check_hip(hipMemcpyAsync(u_aos, g->stage, nv * nc * sizeof(double), hipMemcpyDeviceToHost, g->s), "download u");
    if (!rc) rc = check_hip(hipStreamSynchronize(g->s), "sync");
  }
  if (!rc && flags_aos) {
    int *tmp = (int *)g->stage;
    k_flags_c2i<<<blocks_for(nc * NFLAGS), BS, 0, g->s>>>(nc, g->fl, tmp);
    rc = check_hip(hipMemcpyAsync(flags_aos, tmp, NFLAGS * nc * sizeof(int), hipMemcpyDeviceToHost, g->s), "download flags");
    if (!rc) rc = check_hip(hipStreamSynchronize(g->s), "sync");
  }
  return rc;
}

int koral_hip_grid_device_state(koral_hip_grid *g, double **p, double **u, signed char **flags)
{
  if (!g) return set_error(KORAL_E_ARG, "null grid");
  if (p) *p = g->p;
  if (u) *u = g->u;
  if (flags) *flags = g->fl;
  return 0;
}

int koral_hip_grid_p2u(koral_hip_grid *g)
{
  if (!g) return set_error(KORAL_E_ARG, "null grid");
  const DGrid &G = g->G;
  // u = p2u(p) in every cell (set_initial_profile)
  Box b{-G.ngx, -G.ngy, -G.ngz, G.sx, G.sy, G.sz};
  DISPATCH(g->P, k_p2u_box, <<<blocks_for(b.size()), BS, 0, g->s>>>(g->P, G, g->p, g->u, b));
  return check_launch("grid_p2u");
}

int koral_hip_set_bc(koral_hip_grid *g)
{
  if (!g) return set_error(KORAL_E_ARG, "null grid");
  return run_set_bc(g);
}

int koral_hip_tstepden(koral_hip_grid *g, double *tstepdenmax)
{
  if (!g) return set_error(KORAL_E_ARG, "null grid");
  const DGrid &G = g->G;
  const koral_params &P = g->P;
  int rc = check_hip(hipMemsetAsync(g->tsd, 0, sizeof(double), g->s), "memset tsd");
  if (rc) return rc;
  int lim = (P.int_order == 4) ? 2 : 1;
  int lx = lim, ly = G.ny > 1 ? lim : 0, lz = G.nz > 1 ? lim : 0;
  Box wb{-lx, -ly, -lz, G.nx + 2 * lx, G.ny + 2 * ly, G.nz + 2 * lz};
  DISPATCH(P, k_wavespeeds, <<<blocks_for(wb.size()), BS, 0, g->s>>>(P, G, g->p, g->ws, g->tsd, wb));
  rc = check_hip(hipMemcpyAsync(tstepdenmax, g->tsd, sizeof(double), hipMemcpyDeviceToHost, g->s), "tsd");
  if (!rc) rc = check_hip(hipStreamSynchronize(g->s), "sync");
  return rc;
}

int koral_hip_op_explicit(koral_hip_grid *g, double t, double dt, double *tstepdenmax)
{
  if (!g) return set_error(KORAL_E_ARG, "null grid");
  int rc = check_hip(hipMemsetAsync(g->tsd, 0, sizeof(double), g->s), "memset tsd");
  if (!rc) rc = run_explicit(g, t, dt);
  if (!rc && tstepdenmax) {
    rc = check_hip(hipMemcpyAsync(tstepdenmax, g->tsd, sizeof(double), hipMemcpyDeviceToHost, g->s), "tsd");
    if (!rc) rc = check_hip(hipStreamSynchronize(g->s), "sync");
  }
  return rc;
}

int koral_hip_calc_u2p(koral_hip_grid *g)
{
  if (!g) return set_error(KORAL_E_ARG, "null grid");
  return run_u2p(g);
}

int koral_hip_op_implicit(koral_hip_grid *g, double t, double dt)
{
  if (!g) return set_error(KORAL_E_ARG, "null grid");
  return run_implicit(g, t, dt);
}

int koral_hip_step_rk2imex(koral_hip_grid *g, double t, double dt, double *tstepdenmax, int *counters)
{
  if (!g) return set_error(KORAL_E_ARG, "null grid");
  int rc = check_hip(hipMemsetAsync(g->tsd, 0, sizeof(double), g->s), "memset tsd");
  if (!rc) rc = check_hip(hipMemsetAsync(g->counters, 0, 8 * sizeof(int), g->s), "memset counters");
  if (!rc) rc = run_step(g, t, dt);
  if (!rc && tstepdenmax) rc = check_hip(hipMemcpyAsync(tstepdenmax, g->tsd, sizeof(double), hipMemcpyDeviceToHost, g->s), "tsd");
  if (!rc && counters) rc = check_hip(hipMemcpyAsync(counters, g->counters, 6 * sizeof(int), hipMemcpyDeviceToHost, g->s), "counters");
  if (!rc && (tstepdenmax || counters)) rc = check_hip(hipStreamSynchronize(g->s), "sync");
  return rc;
}

}  // extern "C"

